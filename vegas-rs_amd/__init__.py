"""Python front end of libvegas_hip.so (the C ABI in include/vegas_hip.h).

This is plumbing for Python callers (tests, bench.py): every call goes
through the C ABI into the HIP kernels; there is no CPU path.  Importing it
raises if the library has not been built.

The directory name is not a Python identifier, so load it by path:

    spec = importlib.util.spec_from_file_location(
        "vegas_hip", "vegas-rs_amd/__init__.py")
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libvegas_hip.so"

# lattice kinds, pbc bits, terms, spins, proposals (vegas_hip.h)
SC, BCC = 0, 1
PBC_X, PBC_Y, PBC_Z = 1, 2, 4
EXCHANGE, ZEEMAN, ANISOTROPY, GAUGE = 1, 2, 4, 8
ISING, ISING_BITS, HEISENBERG = 0, 1, 2
FLIP, RANDOM = 0, 1

STATUS = {
    0: "OK", 1: "TemperatureMaxLessThanMin", 2: "NoSteps", 3: "ZeroTemperature",
    4: "ZeroCoolRate", 5: "ZeroField", 6: "ZeroFieldStep", 10: "NotImplemented",
    20: "InvalidArgument", 21: "UnsupportedLattice", 30: "Device", 31: "OutOfMemory",
    32: "NoDevice", 40: "Admission", 41: "NotFound", 42: "Cancelled",
}


class VegasError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code
        self.kind = STATUS.get(code, str(code))


class Lattice(C.Structure):
    _fields_ = [("kind", C.c_int32), ("lx", C.c_uint32), ("ly", C.c_uint32),
                ("lz", C.c_uint32), ("pbc", C.c_uint32)]


class Hamiltonian(C.Structure):
    _fields_ = [("terms", C.c_uint32), ("compound", C.c_int32), ("exchange", C.c_double),
                ("anisotropy", C.c_double), ("anisotropy_axis", C.c_double * 3),
                ("gauge", C.c_double)]


class ChainDesc(C.Structure):
    _fields_ = [("lattice", Lattice), ("spin", C.c_int32), ("proposal", C.c_int32),
                ("hamiltonian", Hamiltonian), ("seed", C.c_uint64)]


class SweepSums(C.Structure):
    _fields_ = [("bond", C.c_int64), ("mag", C.c_int64 * 3), ("axis2", C.c_int64)]


if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is not built (run __graft_entry__.build())")

_lib = C.CDLL(str(LIB_PATH))
_P = C.POINTER
_vp = C.c_void_p
_SIGS = {
    "vg_device_count": [_P(C.c_int)],
    "vg_ctx_create": [C.c_int, _P(_vp)],
    "vg_ctx_destroy": [_vp],
    "vg_batch_create": [_vp, C.c_uint32, _P(ChainDesc), _P(_vp)],
    "vg_batch_destroy": [_vp],
    "vg_batch_size": [_vp, _P(C.c_uint32)],
    "vg_batch_num_sites": [_vp, C.c_uint32, _P(C.c_uint64)],
    "vg_batch_set_thermostat": [_vp, C.c_uint32, C.c_double, C.c_double, _P(C.c_double)],
    "vg_batch_randomize": [_vp],
    "vg_batch_set_ising": [_vp, C.c_uint32, _vp],
    "vg_batch_get_ising": [_vp, C.c_uint32, _vp],
    "vg_batch_set_heisenberg": [_vp, C.c_uint32, _vp],
    "vg_batch_get_heisenberg": [_vp, C.c_uint32, _vp],
    "vg_batch_sweep_count": [_vp, C.c_uint32, _P(C.c_uint64)],
    "vg_batch_set_sweep_count": [_vp, C.c_uint32, C.c_uint64],
    "vg_batch_run": [_vp, C.c_uint64, _vp, _vp],
    "vg_batch_run_sums": [_vp, C.c_uint64, _vp],
    "vg_batch_set_graph": [_vp, C.c_int],
    "vg_batch_last_run_ms": [_vp, _P(C.c_double)],
    "vg_batch_set_timing": [_vp, C.c_int],
    "vg_batch_set_tile": [_vp, C.c_uint32],
    "vg_batch_get_tile": [_vp, _P(C.c_uint32)],
    "vg_batch_autotune": [_vp, C.c_uint32, _P(C.c_uint32)],
    "vg_batch_pass_stats": [_vp, C.c_int32, _P(C.c_double), _P(C.c_uint64)],
}
for _name, _args in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.argtypes = _args
    _f.restype = C.c_int
_lib.vg_last_error.restype = C.c_char_p
_lib.vg_version.restype = C.c_char_p
_lib.vg_sums_scale.argtypes = [C.c_int32]
_lib.vg_sums_scale.restype = C.c_double
_lib.vg_philox4x32_10.argtypes = [_P(C.c_uint32), _P(C.c_uint32), _P(C.c_uint32)]
_lib.vg_philox4x32_10.restype = None


def library() -> C.CDLL:
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise VegasError(rc, _lib.vg_last_error().decode())


def version() -> str:
    return _lib.vg_version().decode()


def device_count() -> int:
    n = C.c_int()
    _check(_lib.vg_device_count(C.byref(n)))
    return n.value


def philox(ctr, key) -> tuple:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    _lib.vg_philox4x32_10(c, k, o)
    return tuple(o)


def sums_scale(spin: int) -> float:
    return float(_lib.vg_sums_scale(spin))


def lattice(kind=SC, lx=4, ly=4, lz=4, pbc=(True, True, True)) -> Lattice:
    bits = (PBC_X if pbc[0] else 0) | (PBC_Y if pbc[1] else 0) | (PBC_Z if pbc[2] else 0)
    return Lattice(kind, lx, ly, lz, bits)


def hamiltonian(exchange=None, zeeman=False, anisotropy=None, axis=(0.0, 0.0, 1.0),
                gauge=None) -> Hamiltonian:
    """Terms present are the ones given; two or more form a Compound."""
    terms = 0
    if exchange is not None:
        terms |= EXCHANGE
    if zeeman:
        terms |= ZEEMAN
    if anisotropy is not None:
        terms |= ANISOTROPY
    if gauge is not None:
        terms |= GAUGE
    compound = 1 if bin(terms).count("1") > 1 else 0
    return Hamiltonian(terms, compound, exchange or 0.0, anisotropy or 0.0,
                       (C.c_double * 3)(*axis), gauge or 0.0)


def chain(lattice: Lattice, spin=ISING, proposal=FLIP, hamiltonian: Hamiltonian | None = None,
          seed=0) -> ChainDesc:
    if hamiltonian is None:
        hamiltonian = globals()["hamiltonian"](exchange=1.0, zeeman=True)
    return ChainDesc(lattice, spin, proposal, hamiltonian, seed)


class Context:
    def __init__(self, device: int = 0):
        h = _vp()
        _check(_lib.vg_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _lib.vg_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Independent chains resident on the device (one Machine each)."""

    def __init__(self, ctx: Context, descs):
        descs = list(descs)
        arr = (ChainDesc * len(descs))(*descs)
        h = _vp()
        _check(_lib.vg_batch_create(ctx.h, len(descs), arr, C.byref(h)))
        self.h = h
        self.ctx = ctx
        self.descs = descs
        self.n = len(descs)

    def close(self):
        if self.h:
            _lib.vg_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def num_sites(self, i: int) -> int:
        n = C.c_uint64()
        _check(_lib.vg_batch_num_sites(self.h, i, C.byref(n)))
        return n.value

    def set_thermostat(self, i: int, temperature: float, field: float = 0.0,
                       field_dir=(0.0, 0.0, 1.0)) -> None:
        d = (C.c_double * 3)(*field_dir)
        _check(_lib.vg_batch_set_thermostat(self.h, i, temperature, field, d))

    def randomize(self) -> None:
        _check(_lib.vg_batch_randomize(self.h))

    def set_ising(self, i: int, spins) -> None:
        a = np.ascontiguousarray(spins, dtype=np.int8)
        _check(_lib.vg_batch_set_ising(self.h, i, a.ctypes.data))

    def get_ising(self, i: int) -> np.ndarray:
        a = np.zeros(self.num_sites(i), dtype=np.int8)
        _check(_lib.vg_batch_get_ising(self.h, i, a.ctypes.data))
        return a

    def set_heisenberg(self, i: int, xyz) -> None:
        a = np.ascontiguousarray(xyz, dtype=np.float64)
        _check(_lib.vg_batch_set_heisenberg(self.h, i, a.ctypes.data))

    def get_heisenberg(self, i: int) -> np.ndarray:
        a = np.zeros((self.num_sites(i), 3), dtype=np.float64)
        _check(_lib.vg_batch_get_heisenberg(self.h, i, a.ctypes.data))
        return a

    def sweep_count(self, i: int) -> int:
        n = C.c_uint64()
        _check(_lib.vg_batch_sweep_count(self.h, i, C.byref(n)))
        return n.value

    def set_sweep_count(self, i: int, sweeps: int) -> None:
        _check(_lib.vg_batch_set_sweep_count(self.h, i, sweeps))

    def run(self, sweeps: int, observe: bool = True):
        """Runs `sweeps` sweeps; returns (energy, |M|) of shape (chains, sweeps)."""
        if not observe:
            _check(_lib.vg_batch_run(self.h, sweeps, None, None))
            return None, None
        e = np.zeros((self.n, sweeps))
        m = np.zeros((self.n, sweeps))
        _check(_lib.vg_batch_run(self.h, sweeps, e.ctypes.data, m.ctypes.data))
        return e, m

    def run_sums(self, sweeps: int) -> np.ndarray:
        """Raw per-sweep sums, int64 array (chains, sweeps, 5):
        bond, mx, my, mz, axis2."""
        a = np.zeros((self.n, sweeps, 5), dtype=np.int64)
        _check(_lib.vg_batch_run_sums(self.h, sweeps, a.ctypes.data))
        return a

    def set_graph(self, enabled: bool) -> None:
        _check(_lib.vg_batch_set_graph(self.h, 1 if enabled else 0))

    def set_tile(self, tile: int) -> None:
        _check(_lib.vg_batch_set_tile(self.h, tile))

    def tile(self) -> int:
        v = C.c_uint32()
        _check(_lib.vg_batch_get_tile(self.h, C.byref(v)))
        return v.value

    def autotune(self, trial_sweeps: int = 20) -> int:
        """Picks the fastest tile by timing; spins and counters are restored."""
        v = C.c_uint32()
        _check(_lib.vg_batch_autotune(self.h, trial_sweeps, C.byref(v)))
        return v.value

    def set_timing(self, enabled: bool) -> None:
        _check(_lib.vg_batch_set_timing(self.h, 1 if enabled else 0))

    def pass_stats(self, spin: int = ISING):
        """(total kernel ms, launches) of the colour-pass kernel."""
        ms = C.c_double()
        n = C.c_uint64()
        _check(_lib.vg_batch_pass_stats(self.h, spin, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def last_run_ms(self) -> float:
        v = C.c_double()
        _check(_lib.vg_batch_last_run_ms(self.h, C.byref(v)))
        return v.value
