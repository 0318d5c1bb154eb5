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
SC, BCC, FCC, GRAPH = 0, 1, 2, 3
PBC_X, PBC_Y, PBC_Z = 1, 2, 4
EXCHANGE, ZEEMAN, ANISOTROPY, GAUGE = 1, 2, 4, 8
ISING, ISING_BITS, HEISENBERG = 0, 1, 2
FLIP, RANDOM, WOLFF = 0, 1, 2

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


class Csr(C.Structure):
    _fields_ = [("n_sites", C.c_uint64), ("row_ptr", C.c_void_p), ("col", C.c_void_p),
                ("value", C.c_void_p)]


class ServerConfig(C.Structure):
    _fields_ = [("memory_bytes", C.c_uint64), ("max_jobs", C.c_uint32),
                ("quantum_sweeps", C.c_uint32), ("tile", C.c_uint32), ("use_graph", C.c_int32),
                ("sweep_budget", C.c_uint64)]


class Stage(C.Structure):
    _fields_ = [("temperature", C.c_double), ("field", C.c_double),
                ("field_dir", C.c_double * 3), ("relax", C.c_uint64), ("measure", C.c_uint64)]


class JobDesc(C.Structure):
    _fields_ = [("chain", ChainDesc), ("n_stages", C.c_uint32), ("stages", C.POINTER(Stage))]


class StageResult(C.Structure):
    _fields_ = [("temperature", C.c_double), ("field", C.c_double), ("energy", C.c_double),
                ("specific_heat", C.c_double), ("magnetization", C.c_double),
                ("susceptibility", C.c_double), ("binder", C.c_double),
                ("measured", C.c_uint64)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("stage", C.c_uint32), ("sweeps_done", C.c_uint64),
                ("sweeps_total", C.c_uint64), ("submit_s", C.c_double), ("start_s", C.c_double),
                ("end_s", C.c_double), ("last_energy", C.c_double),
                ("last_magnetization", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("running", C.c_uint64),
                ("queued", C.c_uint64), ("attempts", C.c_uint64), ("quanta", C.c_uint64),
                ("device_ms", C.c_double), ("pool_bytes", C.c_uint64), ("pool_used", C.c_uint64),
                ("pool_peak", C.c_uint64), ("peak_jobs", C.c_uint32), ("pad", C.c_uint32)]


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = 0, 1, 2, 3, 4


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
    "vg_batch_create_csr": [_vp, C.c_uint32, _P(ChainDesc), _P(Csr), _P(_vp)],
    "vg_batch_colours": [_vp, C.c_uint32, _P(C.c_uint32)],
    "vg_batch_sums_scales": [_vp, C.c_uint32, _P(C.c_double)],
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
    "vg_batch_prepare": [_vp, C.c_uint64],
    "vg_batch_last_run_ms": [_vp, _P(C.c_double)],
    "vg_batch_set_timing": [_vp, C.c_int],
    "vg_batch_set_tile": [_vp, C.c_uint32],
    "vg_batch_get_tile": [_vp, _P(C.c_uint32)],
    "vg_batch_autotune": [_vp, C.c_uint32, _P(C.c_uint32)],
    "vg_batch_set_tie_capacity": [_vp, C.c_uint32],
    "vg_batch_pass_stats": [_vp, C.c_int32, _P(C.c_double), _P(C.c_uint64)],
    "vg_server_create": [_vp, _P(ServerConfig), _P(_vp)],
    "vg_server_destroy": [_vp],
    "vg_server_submit": [_vp, _P(JobDesc), _P(C.c_uint64)],
    "vg_server_status": [_vp, C.c_uint64, _P(JobStatus)],
    "vg_server_wait": [_vp, C.c_uint64, C.c_double, _P(JobStatus)],
    "vg_server_results": [_vp, C.c_uint64, _P(StageResult), C.c_uint32],
    "vg_server_cancel": [_vp, C.c_uint64],
    "vg_server_release": [_vp, C.c_uint64],
    "vg_server_get_stats": [_vp, _P(ServerStats)],
}
for _name, _args in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.argtypes = _args
    _f.restype = C.c_int
_lib.vg_run_sums_noalloc.argtypes = [_P(ChainDesc), C.c_double, C.c_double, _P(C.c_double),
                                     C.c_uint64, C.c_uint64, _vp, _vp, _vp]
_lib.vg_run_sums_noalloc.restype = C.c_int
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


def run_sums_noalloc(desc: ChainDesc, temperature: float, field: float, field_dir,
                     sweep0: int, sweeps: int, spins_ptr: int, sums_ptr: int,
                     stream: int = 0) -> None:
    """vg_run_sums_noalloc: sweeps of one Ising chain on caller-owned device
    buffers (device addresses as ints), enqueued on `stream`."""
    d = (C.c_double * 3)(*field_dir)
    _check(_lib.vg_run_sums_noalloc(C.byref(desc), temperature, field, d, sweep0, sweeps,
                                    _vp(spins_ptr), _vp(sums_ptr), _vp(stream)))


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


def graph_lattice() -> Lattice:
    """The lattice field of a chain whose exchange is a matrix (Batch(...,
    exchange=[...]))."""
    return Lattice(GRAPH, 0, 0, 0, 0)


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

    def __init__(self, ctx: Context, descs, exchange=None):
        """exchange: for batches with graph chains (lattice kind GRAPH), one
        entry per chain -- a CSR matrix (row_ptr, col, value) for a graph
        chain, None for a lattice chain (Exchange::new(matrix),
        src/energy.rs:170-172)."""
        descs = list(descs)
        arr = (ChainDesc * len(descs))(*descs)
        h = _vp()
        if exchange is None:
            _check(_lib.vg_batch_create(ctx.h, len(descs), arr, C.byref(h)))
        else:
            keep, mats = [], (Csr * len(descs))()
            for i, m in enumerate(exchange):
                if m is None:
                    continue
                rp = np.ascontiguousarray(m[0], dtype=np.uint64)
                cl = np.ascontiguousarray(m[1], dtype=np.uint64)
                vl = np.ascontiguousarray(m[2], dtype=np.float64)
                keep += [rp, cl, vl]
                mats[i] = Csr(len(rp) - 1, rp.ctypes.data, cl.ctypes.data, vl.ctypes.data)
            _check(_lib.vg_batch_create_csr(ctx.h, len(descs), arr, mats, C.byref(h)))
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

    def colours(self, i: int) -> int:
        v = C.c_uint32()
        _check(_lib.vg_batch_colours(self.h, i, C.byref(v)))
        return v.value

    def sums_scales(self, i: int) -> np.ndarray:
        v = (C.c_double * 5)()
        _check(_lib.vg_batch_sums_scales(self.h, i, v))
        return np.array(list(v))

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

    def prepare(self, sweeps: int) -> None:
        """Captures and uploads the graphs a run of `sweeps` sweeps replays
        (graphs enabled), running no sweep."""
        _check(_lib.vg_batch_prepare(self.h, sweeps))

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

    def set_tie_capacity(self, records: int) -> None:
        """Test hook: tie records a wave may queue (0..128; the bcc and fcc
        passes hold at most 64) before it settles further tied rows in place;
        results do not depend on it."""
        _check(_lib.vg_batch_set_tie_capacity(self.h, records))

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


# --------------------------------------------------------------- programs --
# Stage lists of the reference's programs (src/program.rs:97-336) with their
# defaults, validation order and f64 loops.  A stage is (temperature, field,
# relax sweeps, measure sweeps); the field acts along +z with magnitude
# |field| (Field::new(S::up(), field), state.rs:219-221).

def _stage(T, field, relax, measure):
    return Stage(T, field, (C.c_double * 3)(0.0, 0.0, 1.0), relax, measure)


def relax(steps=1000, temperature=3.0, field=0.0):
    """Relax (program.rs:97-124): one relax phase, nothing measured."""
    if steps == 0:
        raise VegasError(2, "there should be at least one step")
    if temperature < np.finfo(float).eps:
        raise VegasError(3, "temperature must be greater than zero")
    return [_stage(temperature, field, steps, 0)]


def cool_down(max_temperature=3.0, min_temperature=0.05, cool_rate=0.1, relax=1000,
              steps=20000, field=0.0):
    """CoolDown (program.rs:181-212)."""
    eps = np.finfo(float).eps
    if max_temperature < min_temperature:
        raise VegasError(1, "maximum temperature must be greater than minimum temperature")
    if steps == 0:
        raise VegasError(2, "there should be at least one step")
    if min_temperature < eps:
        raise VegasError(3, "temperature must be greater than zero")
    if cool_rate < eps:
        raise VegasError(4, "cooling rate must be greater than zero")
    out, t = [], max_temperature
    while True:
        out.append(_stage(t, field, relax, steps))
        t -= cool_rate
        if t < min_temperature:
            break
    return out


def hysteresis_loop(steps=1000, relax=1000, temperature=3.0, max_field=1.0, field_step=0.1):
    """HysteresisLoop (program.rs:281-334): 0 -> +max -> -max -> +max."""
    eps = np.finfo(float).eps
    if steps == 0:
        raise VegasError(2, "there should be at least one step")
    if temperature < eps:
        raise VegasError(3, "temperature must be greater than zero")
    if max_field < eps:
        raise VegasError(5, "maximum field must be greater than zero")
    if field_step < eps:
        raise VegasError(6, "field step must be greater than zero")
    out, h = [], 0.0
    while True:
        out.append(_stage(temperature, h, relax, steps))
        h += field_step
        if h > max_field:
            break
    while True:
        out.append(_stage(temperature, h, relax, steps))
        h -= field_step
        if h < -max_field:
            break
    while True:
        out.append(_stage(temperature, h, relax, steps))
        h += field_step
        if h > max_field:
            break
    return out


class Server:
    """Job server with continuous batching (vg_server_* of vegas_hip.h)."""

    def __init__(self, ctx: Context, memory_bytes=1 << 30, max_jobs=256, quantum_sweeps=100,
                 tile=2, use_graph=False, sweep_budget=0):
        cfg = ServerConfig(memory_bytes, max_jobs, quantum_sweeps, tile, 1 if use_graph else 0,
                           sweep_budget)
        h = _vp()
        _check(_lib.vg_server_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self.n_stages = {}

    def close(self):
        if self.h:
            _lib.vg_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, desc: ChainDesc, stages) -> int:
        stages = list(stages)
        arr = (Stage * len(stages))(*stages)
        job = JobDesc(desc, len(stages), arr)
        jid = C.c_uint64()
        _check(_lib.vg_server_submit(self.h, C.byref(job), C.byref(jid)))
        self.n_stages[jid.value] = len(stages)
        return jid.value

    def status(self, jid: int) -> JobStatus:
        st = JobStatus()
        _check(_lib.vg_server_status(self.h, jid, C.byref(st)))
        return st

    def wait(self, jid: int, timeout: float = -1.0) -> JobStatus:
        st = JobStatus()
        _check(_lib.vg_server_wait(self.h, jid, timeout, C.byref(st)))
        return st

    def results(self, jid: int) -> np.ndarray:
        """Per-stage (T, |h|, <E>, C, <|M|>, chi, U4, measured) as float64 rows."""
        n = self.n_stages[jid]
        arr = (StageResult * n)()
        _check(_lib.vg_server_results(self.h, jid, arr, n))
        return np.array([[r.temperature, r.field, r.energy, r.specific_heat, r.magnetization,
                          r.susceptibility, r.binder, float(r.measured)] for r in arr])

    def cancel(self, jid: int) -> None:
        _check(_lib.vg_server_cancel(self.h, jid))

    def release(self, jid: int) -> None:
        _check(_lib.vg_server_release(self.h, jid))
        self.n_stages.pop(jid, None)

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(_lib.vg_server_get_stats(self.h, C.byref(st)))
        return st
