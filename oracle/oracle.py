"""ctypes front end of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
engine never imports it.  See oracle.h for what the oracle restates and which
reference lines each part follows.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"


class Lattice(C.Structure):
    _fields_ = [("kind", C.c_int32), ("lx", C.c_uint32), ("ly", C.c_uint32),
                ("lz", C.c_uint32), ("pbc", C.c_uint32)]


class Model(C.Structure):
    _fields_ = [("spin", C.c_int32), ("proposal", C.c_int32), ("terms", C.c_uint32),
                ("compound", C.c_int32), ("exchange", C.c_double), ("anisotropy", C.c_double),
                ("axis", C.c_double * 3), ("gauge", C.c_double), ("temperature", C.c_double),
                ("field", C.c_double), ("field_dir", C.c_double * 3)]


class Pcg64(C.Structure):
    _fields_ = [("state_hi", C.c_uint64), ("state_lo", C.c_uint64),
                ("inc_hi", C.c_uint64), ("inc_lo", C.c_uint64)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        L.or_philox.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.or_num_sites.argtypes = [P(Lattice)]
        L.or_num_sites.restype = C.c_uint64
        L.or_coordination.argtypes = [P(Lattice)]
        L.or_neighbours.argtypes = [P(Lattice), C.c_uint64, P(C.c_int64)]
        L.or_place.argtypes = [P(Lattice), C.c_uint64, P(C.c_uint32), P(C.c_uint32),
                               P(C.c_uint32)]
        L.or_ising_table.argtypes = [P(Lattice), P(Model), P(C.c_uint32), P(C.c_uint32)]
        L.or_ising_init.argtypes = [P(Lattice), C.c_uint64, C.c_void_p]
        L.or_ising_sweeps.argtypes = [P(Lattice), P(Model), C.c_uint64, C.c_uint64, C.c_uint64,
                                      C.c_void_p, C.c_void_p]
        L.or_heis_init.argtypes = [P(Lattice), C.c_uint64, C.c_void_p]
        L.or_heis_sweeps.argtypes = [P(Lattice), P(Model), C.c_uint64, C.c_uint64, C.c_uint64,
                                     C.c_void_p, C.c_void_p]
        L.or_ising_sweeps_sc_fast.argtypes = [P(Lattice), P(Model), C.c_uint64, C.c_uint64,
                                              C.c_uint64, C.c_void_p, C.c_void_p, C.c_int]
        L.or_wolff_threshold.argtypes = [P(Model), P(C.c_uint32), P(C.c_int)]
        L.or_wolff_steps.argtypes = [P(Lattice), P(Model), C.c_uint64, C.c_uint64, C.c_uint64,
                                     C.c_void_p, C.c_void_p]
        L.or_exp_nonpos.argtypes = [C.c_float]
        L.or_exp_nonpos.restype = C.c_float
        L.or_graph_create.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_graph_create.restype = C.c_void_p
        L.or_graph_destroy.argtypes = [C.c_void_p]
        L.or_graph_colours.argtypes = [C.c_void_p]
        L.or_graph_colours.restype = C.c_uint32
        L.or_graph_colouring.argtypes = [C.c_void_p, C.c_void_p]
        L.or_graph_diag.argtypes = [C.c_void_p]
        L.or_graph_diag.restype = C.c_double
        L.or_exp_f64_nonpos.argtypes = [C.c_double]
        L.or_exp_f64_nonpos.restype = C.c_double
        L.or_graph_ising_init.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.or_graph_ising_sweeps.argtypes = [C.c_void_p, P(Model), C.c_uint64, C.c_uint64,
                                            C.c_uint64, C.c_void_p, C.c_void_p]
        L.or_graph_heis_init.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.or_graph_heis_sweeps.argtypes = [C.c_void_p, P(Model), C.c_uint64, C.c_uint64,
                                           C.c_uint64, C.c_void_p, C.c_void_p]
        L.or_seq_create_csr.argtypes = [P(Model), C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_uint64]
        L.or_seq_create_csr.restype = C.c_void_p
        L.or_pcg64_seed.argtypes = [P(Pcg64), C.c_uint64]
        L.or_pcg64_next.argtypes = [P(Pcg64)]
        L.or_pcg64_next.restype = C.c_uint64
        L.or_seq_create.argtypes = [P(Lattice), P(Model), C.c_uint64]
        L.or_seq_create.restype = C.c_void_p
        L.or_seq_destroy.argtypes = [C.c_void_p]
        L.or_seq_set_thermostat.argtypes = [C.c_void_p, C.c_double, C.c_double, P(C.c_double)]
        L.or_seq_set_state.argtypes = [C.c_void_p, C.c_void_p]
        L.or_seq_get_state.argtypes = [C.c_void_p, C.c_void_p]
        L.or_seq_sweeps.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
        L.or_seq_total_energy.argtypes = [C.c_void_p]
        L.or_seq_total_energy.restype = C.c_double
        L.or_seq_magnetization.argtypes = [C.c_void_p]
        L.or_seq_magnetization.restype = C.c_double
        L.or_seq_nnz.argtypes = [C.c_void_p]
        L.or_seq_nnz.restype = C.c_uint64
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


THR_STRIDE = 25  # classes per current bit: S in [0, 2z], z <= 12


def lattice(kind="sc", lx=4, ly=4, lz=4, pbc=(True, True, True)) -> Lattice:
    bits = (1 if pbc[0] else 0) | (2 if pbc[1] else 0) | (4 if pbc[2] else 0)
    return Lattice({"sc": 0, "bcc": 1, "fcc": 2}[kind], lx, ly, lz, bits)


def model(spin="ising", proposal="flip", terms=("exchange", "zeeman"), exchange=1.0,
          anisotropy=0.0, axis=(0.0, 0.0, 1.0), gauge=0.0, temperature=1.0, field=0.0,
          field_dir=(0.0, 0.0, 1.0)) -> Model:
    bits = 0
    for t in terms:
        bits |= {"exchange": 1, "zeeman": 2, "anisotropy": 4, "gauge": 8}[t]
    return Model(0 if spin == "ising" else 2, 0 if proposal == "flip" else 1, bits,
                 1 if len(terms) > 1 else 0, exchange, anisotropy, (C.c_double * 3)(*axis),
                 gauge, temperature, field, (C.c_double * 3)(*field_dir))


def philox(ctr, key) -> tuple:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox(c, k, o)
    return tuple(o)


def num_sites(L: Lattice) -> int:
    return int(lib().or_num_sites(C.byref(L)))


def neighbours(L: Lattice, site: int) -> list:
    z = lib().or_coordination(C.byref(L))
    out = (C.c_int64 * 12)()
    lib().or_neighbours(C.byref(L), site, out)
    return list(out)[:z]


def place(L: Lattice, site: int) -> tuple:
    c, r, p = C.c_uint32(), C.c_uint32(), C.c_uint32()
    lib().or_place(C.byref(L), site, C.byref(c), C.byref(r), C.byref(p))
    return c.value, r.value, p.value


def ising_table(L: Lattice, M: Model):
    thr = (C.c_uint32 * (2 * THR_STRIDE))()
    alw = (C.c_uint32 * 2)()
    lib().or_ising_table(C.byref(L), C.byref(M), thr, alw)
    return np.array(list(thr), dtype=np.uint64).reshape(2, THR_STRIDE), list(alw)


def ising_init(L: Lattice, seed: int) -> np.ndarray:
    s = np.zeros(num_sites(L), dtype=np.int8)
    lib().or_ising_init(C.byref(L), seed, _ptr(s))
    return s


def ising_sweeps(L: Lattice, M: Model, seed: int, sweep0: int, n: int, spins: np.ndarray):
    """Advances `spins` (natural order, int8 +-1) in place; returns sums (n, 5)."""
    assert spins.dtype == np.int8 and spins.flags.c_contiguous
    sums = np.zeros((n, 5), dtype=np.int64)
    lib().or_ising_sweeps(C.byref(L), C.byref(M), seed, sweep0, n, _ptr(spins), _ptr(sums))
    return sums


def ising_sweeps_sc_fast(L: Lattice, M: Model, seed: int, sweep0: int, n: int,
                         spins: np.ndarray, threads: int = 1):
    """or_ising_sweeps for sc lattices, fast and on `threads` OpenMP threads."""
    assert spins.dtype == np.int8 and spins.flags.c_contiguous and L.kind == 0
    sums = np.zeros((n, 5), dtype=np.int64)
    lib().or_ising_sweeps_sc_fast(C.byref(L), C.byref(M), seed, sweep0, n, _ptr(spins),
                                  _ptr(sums), threads)
    return sums


def wolff_threshold(M: Model) -> tuple:
    thr, al = C.c_uint32(), C.c_int()
    lib().or_wolff_threshold(C.byref(M), C.byref(thr), C.byref(al))
    return thr.value, bool(al.value)


def wolff_steps(L: Lattice, M: Model, seed: int, step0: int, n: int, spins: np.ndarray):
    """Wolff single-cluster steps on `spins` (natural order, int8 +-1) in
    place; returns sums (n, 5) after each step."""
    assert spins.dtype == np.int8 and spins.flags.c_contiguous
    sums = np.zeros((n, 5), dtype=np.int64)
    lib().or_wolff_steps(C.byref(L), C.byref(M), seed, step0, n, _ptr(spins), _ptr(sums))
    return sums


def heis_init(L: Lattice, seed: int) -> np.ndarray:
    s = np.zeros((num_sites(L), 3), dtype=np.float32)
    lib().or_heis_init(C.byref(L), seed, _ptr(s))
    return s


def heis_sweeps(L: Lattice, M: Model, seed: int, sweep0: int, n: int, xyz: np.ndarray):
    assert xyz.dtype == np.float32 and xyz.flags.c_contiguous
    sums = np.zeros((n, 5), dtype=np.int64)
    lib().or_heis_sweeps(C.byref(L), C.byref(M), seed, sweep0, n, _ptr(xyz), _ptr(sums))
    return sums


def exp_nonpos(x: float) -> float:
    return float(lib().or_exp_nonpos(x))


def exp_f64_nonpos(x: float) -> float:
    return float(lib().or_exp_f64_nonpos(x))


def csr_arrays(row_ptr, col, val):
    """Contiguous uint64 / uint64 / float64 copies of a CSR matrix."""
    return (np.ascontiguousarray(row_ptr, dtype=np.uint64),
            np.ascontiguousarray(col, dtype=np.uint64),
            np.ascontiguousarray(val, dtype=np.float64))


class Graph:
    """A chain's exchange given as an arbitrary CSR matrix (Exchange::new,
    energy.rs:170-172), swept by the colour classes of its greedy colouring."""

    def __init__(self, row_ptr, col, val):
        self.row_ptr, self.col, self.val = csr_arrays(row_ptr, col, val)
        self.n = len(self.row_ptr) - 1
        self.h = lib().or_graph_create(self.n, _ptr(self.row_ptr), _ptr(self.col),
                                       _ptr(self.val))

    def close(self):
        if self.h:
            lib().or_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def colours(self) -> int:
        return int(lib().or_graph_colours(self.h))

    def colouring(self) -> np.ndarray:
        c = np.zeros(self.n, dtype=np.uint32)
        lib().or_graph_colouring(self.h, _ptr(c))
        return c

    def diag(self) -> float:
        return float(lib().or_graph_diag(self.h))

    def ising_init(self, seed: int) -> np.ndarray:
        s = np.zeros(self.n, dtype=np.int8)
        lib().or_graph_ising_init(self.h, seed, _ptr(s))
        return s

    def ising_sweeps(self, M: Model, seed: int, sweep0: int, n: int, spins: np.ndarray):
        assert spins.dtype == np.int8 and spins.flags.c_contiguous and len(spins) == self.n
        sums = np.zeros((n, 5), dtype=np.int64)
        lib().or_graph_ising_sweeps(self.h, C.byref(M), seed, sweep0, n, _ptr(spins), _ptr(sums))
        return sums

    def heis_init(self, seed: int) -> np.ndarray:
        s = np.zeros((self.n, 3), dtype=np.float32)
        lib().or_graph_heis_init(self.h, seed, _ptr(s))
        return s

    def heis_sweeps(self, M: Model, seed: int, sweep0: int, n: int, xyz: np.ndarray):
        assert xyz.dtype == np.float32 and xyz.flags.c_contiguous and len(xyz) == self.n
        sums = np.zeros((n, 5), dtype=np.int64)
        lib().or_graph_heis_sweeps(self.h, C.byref(M), seed, sweep0, n, _ptr(xyz), _ptr(sums))
        return sums


class Sequential:
    """The reference's random-site Metropolis (integrator.rs:66-138) in f64,
    on a lattice or (csr=(row_ptr, col, val)) on a caller's exchange matrix."""

    def __init__(self, L, M: Model, seed: int, csr=None):
        self.L, self.M = L, M
        if csr is not None:
            self._csr = csr_arrays(*csr)
            self.n = len(self._csr[0]) - 1
            self.h = lib().or_seq_create_csr(C.byref(M), self.n, _ptr(self._csr[0]),
                                             _ptr(self._csr[1]), _ptr(self._csr[2]), seed)
            return
        self.n = num_sites(L)
        self.h = lib().or_seq_create(C.byref(L), C.byref(M), seed)

    def close(self):
        if self.h:
            lib().or_seq_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_thermostat(self, temperature, field=0.0, field_dir=(0.0, 0.0, 1.0)):
        d = (C.c_double * 3)(*field_dir)
        lib().or_seq_set_thermostat(self.h, temperature, field, d)

    def set_state(self, spins: np.ndarray):
        a = np.ascontiguousarray(spins)
        lib().or_seq_set_state(self.h, _ptr(a))

    def get_state(self) -> np.ndarray:
        if self.M.spin == 0:
            a = np.zeros(self.n, dtype=np.int8)
        else:
            a = np.zeros((self.n, 3), dtype=np.float64)
        lib().or_seq_get_state(self.h, _ptr(a))
        return a

    def sweeps(self, n: int, observe: bool = True):
        if not observe:
            lib().or_seq_sweeps(self.h, n, None, None)
            return None, None
        e = np.zeros(n)
        m = np.zeros(n)
        lib().or_seq_sweeps(self.h, n, _ptr(e), _ptr(m))
        return e, m

    def total_energy(self) -> float:
        return float(lib().or_seq_total_energy(self.h))

    def magnetization(self) -> float:
        return float(lib().or_seq_magnetization(self.h))

    def nnz(self) -> int:
        return int(lib().or_seq_nnz(self.h))


def pcg64(seed: int) -> Pcg64:
    g = Pcg64()
    lib().or_pcg64_seed(C.byref(g), seed)
    return g


def pcg64_next(g: Pcg64) -> int:
    return int(lib().or_pcg64_next(C.byref(g)))

