"""Pins of the CPU oracle that do not need a GPU.

Each check is chosen so that a plausible slip in the oracle (a wrong sign,
a dropped term, a transposed index, a wrong neighbour) fails it: known
answers from independent implementations (numpy's PCG64), closed forms,
exact enumeration of tiny lattices, and the reference's own unit tests
(energy.rs:300-375, state.rs:330-382) with two-sided expectations.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from _stats import binned, ising_exact, z_score


# ---------------------------------------------------------------- PCG64 ----
def test_pcg64_matches_numpy(orc):
    """The sequential oracle's generator is PCG64 XSL-RR 128/64, the algorithm
    of rand_pcg::Lcg128Xsl64 (Cargo.toml:14) and numpy's PCG64."""
    g = orc.pcg64(12345)
    state = (g.state_hi << 64) | g.state_lo
    inc = (g.inc_hi << 64) | g.inc_lo
    bg = np.random.PCG64()
    bg.state = {"bit_generator": "PCG64", "state": {"state": state, "inc": inc},
                "has_uint32": 0, "uinteger": 0}
    expected = bg.random_raw(1000)
    got = [orc.pcg64_next(g) for _ in range(1000)]
    assert got == [int(v) for v in expected]


# -------------------------------------------------------------- lattice ----
@pytest.mark.parametrize("kind,dims,pbc", [
    ("sc", (4, 4, 4), (1, 1, 1)), ("sc", (10, 10, 1), (1, 1, 0)), ("sc", (2, 4, 6), (1, 1, 1)),
    ("sc", (6, 3, 5), (1, 0, 0)), ("bcc", (3, 3, 3), (1, 1, 1)), ("bcc", (2, 3, 1), (0, 1, 0)),
    ("fcc", (3, 3, 3), (1, 1, 1)), ("fcc", (3, 2, 4), (0, 1, 1)), ("fcc", (2, 3, 1), (1, 0, 0)),
])
def test_neighbours_symmetric_and_coloured(orc, kind, dims, pbc):
    L = orc.lattice(kind, *dims, pbc=pbc)
    n = orc.num_sites(L)
    counts = {}
    for i in range(n):
        ci, _, _ = orc.place(L, i)
        for j in orc.neighbours(L, i):
            if j < 0:
                continue
            assert orc.place(L, j)[0] != ci  # the colours are independent sets
            counts[(i, j)] = counts.get((i, j), 0) + 1
    for (i, j), c in counts.items():
        assert counts.get((j, i)) == c  # symmetric with multiplicity


def test_neighbour_counts(orc):
    # sc periodic: 6 distinct neighbours; open 10x10x1: 4 in the bulk, 0 along z
    L = orc.lattice("sc", 4, 4, 4)
    assert all(len([j for j in orc.neighbours(L, i) if j >= 0]) == 6 for i in range(64))
    L = orc.lattice("sc", 10, 10, 1, pbc=(1, 1, 0))
    for i in range(100):
        nb = orc.neighbours(L, i)
        assert nb[4] == -1 and nb[5] == -1 and min(nb[:4]) >= 0
    # site (1,2,3) of a 4x5x6 sc lattice: explicit neighbours
    L = orc.lattice("sc", 4, 5, 6)
    site = 1 + 4 * (2 + 5 * 3)
    expect = [0 + 4 * (2 + 5 * 3), 2 + 4 * (2 + 5 * 3), 1 + 4 * (1 + 5 * 3), 1 + 4 * (3 + 5 * 3),
              1 + 4 * (2 + 5 * 2), 1 + 4 * (2 + 5 * 4)]
    assert orc.neighbours(L, site) == expect
    # bcc periodic: 8 distinct neighbours on the other sublattice
    L = orc.lattice("bcc", 3, 3, 3)
    for i in range(orc.num_sites(L)):
        nb = orc.neighbours(L, i)
        assert len(set(nb)) == 8 and all((j & 1) != (i & 1) for j in nb)
    # fcc periodic: 12 distinct neighbours, 4 on each other sublattice
    L = orc.lattice("fcc", 3, 3, 3)
    for i in range(orc.num_sites(L)):
        nb = orc.neighbours(L, i)
        assert len(set(nb)) == 12
        assert sorted(j & 3 for j in nb) == sorted([b for b in range(4) if b != (i & 3)] * 4)


FCC_HALF = [(0, 0, 0), (0, 1, 1), (1, 0, 1), (1, 1, 0)]


@pytest.mark.parametrize("dims,pbc", [((3, 3, 3), (1, 1, 1)), ((3, 4, 3), (0, 1, 0)),
                                      ((4, 3, 5), (1, 0, 1))])
def test_fcc_neighbours_are_the_geometric_ones(orc, dims, pbc):
    """fcc neighbours from positions: the sites at distance a/sqrt(2) (the
    nearest-neighbour distance of the fcc cell), by brute force over all
    pairs with minimum-image distances along the periodic axes."""
    L = orc.lattice("fcc", *dims, pbc=pbc)
    n = orc.num_sites(L)
    pos = np.zeros((n, 3))
    for i in range(n):
        cell, sub = i >> 2, i & 3
        c = (cell % dims[0], (cell // dims[0]) % dims[1], cell // (dims[0] * dims[1]))
        pos[i] = [c[e] + FCC_HALF[sub][e] / 2 for e in range(3)]
    for i in range(n):
        d = pos - pos[i]
        for e in range(3):
            if pbc[e]:
                d[:, e] -= dims[e] * np.round(d[:, e] / dims[e])
        geo = set(np.nonzero(np.abs((d * d).sum(axis=1) - 0.5) < 1e-9)[0].tolist())
        nb = [j for j in orc.neighbours(L, i) if j >= 0]
        assert len(nb) == len(set(nb)) and set(nb) == geo, i


def test_place_round_trip(orc):
    for kind, dims in [("sc", (6, 4, 2)), ("bcc", (3, 2, 2)), ("fcc", (3, 2, 2))]:
        L = orc.lattice(kind, *dims)
        seen = set()
        for i in range(orc.num_sites(L)):
            seen.add(orc.place(L, i))
        assert len(seen) == orc.num_sites(L)


# ---------------------------------------------- reference energy quirks ----
def _heis_seq(orc, n, terms, **kw):
    L = orc.lattice("sc", n, 1, 1, pbc=(0, 0, 0))
    M = orc.model(spin="heisenberg", proposal="random", terms=terms, **kw)
    return orc.Sequential(L, M, 1)


def test_reference_energy_unit_tests(orc):
    """energy.rs:300-375 with the reference's values computed, not its
    one-sided asserts: Gauge(10) -> 100; Anisotropy total omits K -> 10;
    Zeeman total is -|h| sum s.h -> -10/+10 and -20/+20; Compound(Gauge(10),
    Anisotropy(1)) uses the default sum of site energies -> 110;
    hamiltonian!(Anisotropy(1), Gauge(1)) -> 20."""
    up = np.tile([0.0, 0.0, 1.0], (10, 1))
    down = -up
    s = _heis_seq(orc, 10, ("gauge",), gauge=10.0)
    s.set_state(up)
    assert s.total_energy() == pytest.approx(100.0, abs=1e-12)
    s = _heis_seq(orc, 10, ("anisotropy",), anisotropy=1.0)
    for st in (up, down):
        s.set_state(st)
        assert s.total_energy() == pytest.approx(10.0, abs=1e-12)
    s = _heis_seq(orc, 10, ("anisotropy",), anisotropy=7.0)
    s.set_state(up)
    assert s.total_energy() == pytest.approx(10.0, abs=1e-12)  # K omitted (Q3)
    for mag in (1.0, 2.0):
        s = _heis_seq(orc, 10, ("zeeman",))
        s.set_thermostat(1e-300, mag, (0.0, 0.0, 1.0))
        s.set_state(up)
        assert s.total_energy() == pytest.approx(-10.0 * mag, abs=1e-12)
        s.set_state(down)
        assert s.total_energy() == pytest.approx(10.0 * mag, abs=1e-12)
    s = _heis_seq(orc, 10, ("anisotropy", "gauge"), anisotropy=1.0, gauge=10.0)
    s.set_state(up)
    assert s.total_energy() == pytest.approx(110.0, abs=1e-12)
    s = _heis_seq(orc, 10, ("anisotropy", "gauge"), anisotropy=1.0, gauge=1.0)
    s.set_state(up)
    assert s.total_energy() == pytest.approx(20.0, abs=1e-12)


def test_ground_state_energies(orc):
    """Compound(Exchange, Zeeman) double counts exchange (Q4): the 10x10x1
    CLI lattice (x,y periodic, z open) has E/N = -4 all up; Exchange alone on
    sc has E/N = -3 (the `vegas bench` Hamiltonian); bcc E/N = -4."""
    L = orc.lattice("sc", 10, 10, 1, pbc=(1, 1, 0))
    s = orc.Sequential(L, orc.model(terms=("exchange", "zeeman")), 1)
    s.set_state(np.ones(100, dtype=np.int8))
    assert s.total_energy() / 100 == pytest.approx(-4.0)
    assert s.magnetization() == 100.0
    assert s.nnz() == 400
    L = orc.lattice("sc", 6, 6, 6)
    s = orc.Sequential(L, orc.model(terms=("exchange",)), 1)
    s.set_state(-np.ones(216, dtype=np.int8))
    assert s.total_energy() / 216 == pytest.approx(-3.0)
    assert s.magnetization() == 216.0
    L = orc.lattice("bcc", 3, 3, 3)
    s = orc.Sequential(L, orc.model(terms=("exchange",)), 1)
    s.set_state(np.ones(54, dtype=np.int8))
    assert s.total_energy() / 54 == pytest.approx(-4.0)
    L = orc.lattice("fcc", 3, 3, 3)  # 12 bonds per site: E/N = -6
    s = orc.Sequential(L, orc.model(terms=("exchange",)), 1)
    s.set_state(np.ones(108, dtype=np.int8))
    assert s.total_energy() / 108 == pytest.approx(-6.0)
    assert s.nnz() == 108 * 12
    # Neel state: every bond broken
    L = orc.lattice("sc", 4, 4, 4)
    st = np.array([1 if orc.place(L, i)[0] == 0 else -1 for i in range(64)], dtype=np.int8)
    s = orc.Sequential(L, orc.model(terms=("exchange",)), 1)
    s.set_state(st)
    assert s.total_energy() / 64 == pytest.approx(3.0)
    assert s.magnetization() == 0.0


def test_length_two_axis_counts_bond_twice(orc):
    """On a periodic axis of length 2 the +x and -x neighbours coincide; the
    CSR sums the duplicate, so the bond carries 2J."""
    L = orc.lattice("sc", 2, 4, 4)
    s = orc.Sequential(L, orc.model(terms=("exchange",)), 1)
    s.set_state(np.ones(32, dtype=np.int8))
    assert s.total_energy() / 32 == pytest.approx(-3.0)
    assert s.nnz() == 32 * 5


# ---------------------------------------------------- acceptance table ----
def test_ising_table_closed_forms(orc):
    L = orc.lattice("sc", 4, 4, 4)
    T = 3.0
    thr, alw = orc.ising_table(L, orc.model(terms=("exchange",), temperature=T))
    # h = 0: a flip is free or downhill iff at most 3 of 6 neighbours align
    # (S = 2 * up neighbours when none is absent)
    for S in range(0, 13, 2):
        up_aligned = S // 2
        assert bool(alw[1] >> S & 1) == (up_aligned <= 3)
        assert bool(alw[0] >> S & 1) == (6 - up_aligned <= 3)
    # odd S: one absent neighbour slot counts 1, so sum_j s_j = S - 6
    for S in range(1, 13, 2):
        assert bool(alw[1] >> S & 1) == (S - 6 <= 0)
        assert bool(alw[0] >> S & 1) == (6 - S <= 0)
    # all six neighbours aligned: dE = 12 J
    assert thr[1, 12] == math.ceil(math.exp(-12.0 / T) * 2 ** 32)
    assert thr[0, 0] == thr[1, 12]
    assert thr[1, 10] == math.ceil(math.exp(-8.0 / T) * 2 ** 32)
    # Zeeman sign (Q1): the per-site energy +|h| s.h makes Up costly under an
    # Up field, so flipping Up -> Down with no exchange bias is downhill.
    thr, alw = orc.ising_table(L, orc.model(terms=("exchange", "zeeman"), temperature=T,
                                            field=0.5))
    assert alw[1] >> 6 & 1
    assert thr[0, 6] == math.ceil(math.exp(-1.0 / T) * 2 ** 32)
    thr2, alw2 = orc.ising_table(L, orc.model(terms=("exchange", "zeeman"), temperature=T,
                                              field=-0.5))
    assert (thr2 == thr).all() and alw2 == alw  # |magnitude| is used (Q2)
    thr3, alw3 = orc.ising_table(L, orc.model(terms=("exchange", "zeeman"), temperature=T,
                                              field=0.5, field_dir=(0, 0, -1)))
    assert alw3[0] >> 6 & 1 and thr3[1, 6] == thr[0, 6]


# ------------------------------------------------------------- exp(x) ----
def test_exp_nonpos_accuracy(orc):
    xs = np.concatenate([-np.geomspace(1e-8, 86.0, 4000), [0.0, -1e-30, -0.5, -86.0]])
    worst = 0.0
    for x in xs:
        x32 = float(np.float32(x))
        got = orc.exp_nonpos(x32)
        ref = math.exp(x32)
        worst = max(worst, abs(got - ref) / ref)
    assert worst < 3e-7
    assert orc.exp_nonpos(0.0) == 1.0
    assert orc.exp_nonpos(-87.0) == 0.0


# ------------------------------------------------- Heisenberg sampling ----
def test_heisenberg_init_uniform_on_sphere(orc):
    L = orc.lattice("sc", 32, 32, 32)
    s = orc.heis_init(L, 7).astype(np.float64)
    norm = np.linalg.norm(s, axis=1)
    assert np.abs(norm - 1.0).max() < 1e-6
    n = len(s)
    assert np.abs(s.mean(axis=0)).max() < 5 * np.sqrt(1 / 3 / n)
    assert np.abs((s ** 2).mean(axis=0) - 1 / 3).max() < 5 * np.sqrt(4 / 45 / n)
    # the two colour planes are independent streams
    assert not np.array_equal(s[0], s[1])


def test_ising_init_fair(orc):
    L = orc.lattice("sc", 64, 64, 16)
    s = orc.ising_init(L, 42)
    n = len(s)
    assert set(np.unique(s)) == {-1, 1}
    assert abs(s.mean()) < 5 / np.sqrt(n)
    assert not np.array_equal(s, orc.ising_init(L, 43))


# --------------------------------------- checkerboard vs exact and closed ----
def test_checkerboard_ising_matches_exact_enumeration(orc):
    """4x4x1 (x,y periodic, z open) Compound(Exchange, Zeeman): the
    checkerboard oracle's long-run averages equal the exact Boltzmann ones."""
    L = orc.lattice("sc", 4, 4, 1, pbc=(1, 1, 0))
    n = 16
    neigh = [[j for j in orc.neighbours(L, i) if j >= 0] for i in range(n)]
    for T, h in [(2.5, 0.0), (1.8, 0.4), (4.0, -0.7)]:
        M = orc.model(terms=("exchange", "zeeman"), temperature=T, field=h)
        spins = orc.ising_init(L, 3)
        orc.ising_sweeps(L, M, 11, 0, 2000, spins)
        sums = orc.ising_sweeps(L, M, 11, 2000, 200000, spins)
        bond, mag = sums[:, 0].astype(float), sums[:, 3].astype(float)
        e = -2.0 * bond + abs(h) * mag
        ee, em, _, _ = ising_exact(neigh, n, 1.0, h, 1.0, T)
        me, se = binned(e)
        mm, sm = binned(np.abs(mag))
        assert abs(me - ee) < 4 * se, (T, h, me, ee, se)
        assert abs(mm - em) < 4 * sm, (T, h, mm, em, sm)


@pytest.mark.parametrize("dims,pbc,T,h", [((2, 2, 2), (1, 1, 1), 5.0, 0.3),
                                          ((3, 2, 1), (1, 0, 0), 2.0, 0.0)])
def test_checkerboard_bcc_ising_matches_exact_enumeration(orc, dims, pbc, T, h):
    """bcc (sublattices as colours): the checkerboard oracle's long-run
    averages equal the exact Boltzmann ones, with open boundaries too."""
    L = orc.lattice("bcc", *dims, pbc=pbc)
    n = orc.num_sites(L)
    neigh = [[j for j in orc.neighbours(L, i) if j >= 0] for i in range(n)]
    M = orc.model(terms=("exchange", "zeeman"), temperature=T, field=h)
    spins = orc.ising_init(L, 4)
    orc.ising_sweeps(L, M, 13, 0, 2000, spins)
    sums = orc.ising_sweeps(L, M, 13, 2000, 100000, spins)
    bond, mag = sums[:, 0].astype(float), sums[:, 3].astype(float)
    e = -2.0 * bond + abs(h) * mag
    ee, em, _, _ = ising_exact(neigh, n, 1.0, h, 1.0, T)
    me, se = binned(e)
    mm, sm = binned(np.abs(mag))
    assert abs(me - ee) < 4 * se, (me, ee, se)
    assert abs(mm - em) < 4 * sm, (mm, em, sm)


@pytest.mark.parametrize("dims,pbc,T,h", [((2, 1, 1), (1, 1, 1), 6.0, 0.3),
                                          ((2, 2, 1), (1, 1, 0), 4.0, 0.0),
                                          ((1, 2, 2), (0, 1, 0), 2.5, -0.5)])
def test_checkerboard_fcc_ising_matches_exact_enumeration(orc, dims, pbc, T, h):
    """fcc (four sublattice colours): the checkerboard oracle's long-run
    averages equal the exact Boltzmann ones, duplicate bonds of short
    periodic axes and open boundaries included."""
    L = orc.lattice("fcc", *dims, pbc=pbc)
    n = orc.num_sites(L)
    neigh = [[j for j in orc.neighbours(L, i) if j >= 0] for i in range(n)]
    M = orc.model(terms=("exchange", "zeeman"), temperature=T, field=h)
    spins = orc.ising_init(L, 5)
    orc.ising_sweeps(L, M, 17, 0, 2000, spins)
    sums = orc.ising_sweeps(L, M, 17, 2000, 100000, spins)
    bond, mag = sums[:, 0].astype(float), sums[:, 3].astype(float)
    e = -2.0 * bond + abs(h) * mag
    ee, em, _, _ = ising_exact(neigh, n, 1.0, h, 1.0, T)
    me, se = binned(e)
    mm, sm = binned(np.abs(mag))
    assert abs(me - ee) < 4 * se, (me, ee, se)
    assert abs(mm - em) < 4 * sm, (mm, em, sm)


def test_sequential_matches_exact_enumeration(orc):
    L = orc.lattice("sc", 4, 4, 1, pbc=(1, 1, 0))
    n = 16
    neigh = [[j for j in orc.neighbours(L, i) if j >= 0] for i in range(n)]
    for T, h, proposal in [(2.5, 0.0, "flip"), (1.8, 0.4, "random")]:
        M = orc.model(terms=("exchange", "zeeman"), temperature=T, field=h, proposal=proposal)
        s = orc.Sequential(L, M, 5)
        s.set_state(np.ones(n, dtype=np.int8))
        s.set_thermostat(T, h)
        s.sweeps(2000, observe=False)
        e, m = s.sweeps(100000)
        ee, em, _, _ = ising_exact(neigh, n, 1.0, h, 1.0, T)
        me, se = binned(e)
        mm, sm = binned(m)
        assert abs(me - ee) < 4 * se, (T, h, me, ee, se)
        assert abs(mm - em) < 4 * sm, (T, h, mm, em, sm)


def test_checkerboard_ising_random_proposal_matches_exact(orc):
    L = orc.lattice("sc", 4, 4, 1, pbc=(1, 1, 0))
    neigh = [[j for j in orc.neighbours(L, i) if j >= 0] for i in range(16)]
    M = orc.model(terms=("exchange", "zeeman"), temperature=2.2, field=0.3, proposal="random")
    spins = orc.ising_init(L, 8)
    sums = orc.ising_sweeps(L, M, 12, 0, 150000, spins)[1000:]
    e = -2.0 * sums[:, 0] + 0.3 * sums[:, 3]
    ee, em, _, _ = ising_exact(neigh, 16, 1.0, 0.3, 1.0, 2.2)
    me, se = binned(e)
    mm, sm = binned(np.abs(sums[:, 3]))
    assert abs(me - ee) < 4 * se and abs(mm - em) < 4 * sm


def _langevin(a):
    return 1.0 / math.tanh(a) - 1.0 / a


def test_checkerboard_heisenberg_single_spin_closed_forms(orc):
    """With only Zeeman, sites are independent: <s.h> = -L(|h|/T) (the
    per-site energy is +|h| s.h, Q1).  With only anisotropy K along z,
    <s_z^2> = int x^2 e^{-K x^2/T} / int e^{-K x^2/T} over [-1,1]."""
    L = orc.lattice("sc", 8, 8, 8)
    h, T = 0.6, 0.8
    M = orc.model(spin="heisenberg", proposal="random", terms=("zeeman",), temperature=T,
                  field=h)
    xyz = orc.heis_init(L, 1)
    sums = orc.heis_sweeps(L, M, 2, 0, 3000, xyz)[100:]
    mz = sums[:, 3] / 2 ** 27 / 512
    mean, err = binned(mz)
    assert z_score(mean, err, -_langevin(h / T), 0.0) < 4
    K = 1.5
    M = orc.model(spin="heisenberg", proposal="random", terms=("anisotropy",), anisotropy=K,
                  temperature=T)
    sums = orc.heis_sweeps(L, M, 3, 0, 3000, xyz)[100:]
    a2 = sums[:, 4] / 2 ** 27 / 512
    num = integrate.quad(lambda x: x * x * math.exp(-K * x * x / T), -1, 1)[0]
    den = integrate.quad(lambda x: math.exp(-K * x * x / T), -1, 1)[0]
    mean, err = binned(a2)
    assert z_score(mean, err, num / den, 0.0) < 4


def test_ising_10x10_matches_transfer_matrix(orc):
    """The reference's CPU-runnable lattice (sc 10x10x1, x and y periodic,
    J = 1, h = 0; Compound energy): long runs of the checkerboard
    restatement and of the sequential restatement of MetropolisFlipIntegrator
    reproduce the exact <E>/N and <|M|>/N of the 10x10 torus (row transfer
    matrix, tests/golden/make_ising10x10_exact.py) within 4 sigma, at a
    temperature above, near and below T_c = 2.269."""
    import json
    from pathlib import Path
    rows = json.loads((Path(__file__).resolve().parent / "golden" /
                       "ising10x10_exact.json").read_text())["rows"]
    L = orc.lattice("sc", 10, 10, 1, pbc=(1, 1, 0))
    for T in (3.0, 2.3, 1.9):
        ex = min(rows, key=lambda r: abs(r["T"] - T))   # schedule temperatures are f64 sums
        assert abs(ex["T"] - T) < 1e-9
        M = orc.model(terms=("exchange", "zeeman"), temperature=T)
        s = orc.ising_init(L, 3)
        sums = orc.ising_sweeps(L, M, 3, 0, 150000, s)[2000:]
        e, m = -2.0 * sums[:, 0] / 100, np.abs(sums[:, 3]) / 100
        for val, want in ((binned(e), ex["E/N"]), (binned(m), ex["|M|/N"])):
            assert z_score(*val, want, 0.0) < 4, (T, val, want)
        seq = orc.Sequential(L, M, 4)
        seq.set_thermostat(T)
        seq.set_state(orc.ising_init(L, 4))
        seq.sweeps(2000, observe=False)
        e, m = seq.sweeps(150000)
        seq.close()
        for val, want in ((binned(e / 100), ex["E/N"]), (binned(m / 100), ex["|M|/N"])):
            assert z_score(*val, want, 0.0) < 4, ("sequential", T, val, want)


def test_heisenberg_exchange_dimer_closed_form(orc):
    """Two unit spins joined by one exchange bond (sc 2x1x1, open axes):
    with H = -J s1.s2 the bond's distribution is p(c) ~ exp(J c / T) on
    c = s1.s2 in [-1, 1], so <s1.s2> = coth(J/T) - T/J (Langevin L(J/T)).
    Both the checkerboard restatement (bond sum of the fixed-point sums) and
    the sequential restatement of the reference (its Exchange-only
    total_energy is -J s1.s2) must reproduce it; a flipped sign (-L), a
    doubled coupling (L(2J/T)) or a missing bond (0) fail by many sigma."""
    L = orc.lattice("sc", 2, 1, 1, pbc=(0, 0, 0))
    assert orc.neighbours(L, 0) == [-1, 1, -1, -1, -1, -1]
    for J, T in [(1.0, 2.0), (1.0, 1.0), (1.0, 0.5), (-0.7, 0.8)]:
        want = _langevin(J / T)
        M = orc.model(spin="heisenberg", proposal="random", terms=("exchange",), exchange=J,
                      temperature=T)
        xyz = orc.heis_init(L, 7)
        sums = orc.heis_sweeps(L, M, 8, 0, 200000, xyz)[200:]
        c = sums[:, 0] / 2 ** 27
        mean, err = binned(c)
        assert err < 0.01
        assert z_score(mean, err, want, 0.0) < 4, (J, T, mean, err, want)
        for wrong in (-want, _langevin(2 * J / T), 0.0):
            assert z_score(mean, err, wrong, 0.0) > 10, (J, T, wrong)
        seq = orc.Sequential(L, M, 9)
        seq.set_thermostat(T)
        seq.set_state(orc.heis_init(L, 9).astype(np.float64))
        seq.sweeps(200, observe=False)
        e, _ = seq.sweeps(200000)
        seq.close()
        mean, err = binned(-e / J)
        assert z_score(mean, err, want, 0.0) < 4, (J, T, mean, err, want)


def test_checkerboard_ising_zeeman_only_closed_form(orc):
    L = orc.lattice("sc", 8, 8, 8)
    h, T = 0.5, 1.3
    M = orc.model(terms=("zeeman",), temperature=T, field=h)
    spins = orc.ising_init(L, 4)
    sums = orc.ising_sweeps(L, M, 9, 0, 4000, spins)[50:]
    mean, err = binned(sums[:, 3] / 512)
    assert z_score(mean, err, -math.tanh(h / T), 0.0) < 4


def test_heisenberg_flip_proposal_conserves_norm(orc):
    L = orc.lattice("sc", 6, 6, 6)
    M = orc.model(spin="heisenberg", proposal="flip", terms=("exchange", "zeeman"),
                  temperature=1.0, field=0.2)
    xyz = orc.heis_init(L, 5)
    before = np.linalg.norm(xyz, axis=1)
    orc.heis_sweeps(L, M, 6, 0, 20, xyz)
    assert np.array_equal(np.linalg.norm(xyz, axis=1), before)


# ------------------------------------------------------------ Wolff ----
def test_wolff_threshold_closed_form(orc):
    """Bond probability 1 - exp(-2J/T) (integrator.rs:188-192) as a 32-bit
    threshold ceil(p 2^32); J = 0 gives no bonds, T -> 0 bonds always."""
    for T, J in [(2.0, 1.0), (4.5115, 1.0), (0.7, 0.5), (10.0, 2.0)]:
        thr, al = orc.wolff_threshold(orc.model(terms=("exchange",), exchange=J, temperature=T))
        p = 1.0 - math.exp(-2.0 * J / T)
        assert not al and thr == math.ceil(p * 2 ** 32)
    assert orc.wolff_threshold(orc.model(terms=("zeeman",), temperature=1.0)) == (0, False)
    assert orc.wolff_threshold(orc.model(terms=("exchange",), temperature=1e-6))[1]


@pytest.mark.parametrize("kind,dims,pbc,T", [("sc", (4, 4, 1), (1, 1, 0), 2.3),
                                             ("sc", (4, 2, 2), (1, 1, 1), 3.5),
                                             ("bcc", (2, 2, 2), (1, 1, 1), 5.0),
                                             ("fcc", (2, 2, 1), (1, 1, 0), 7.0)])
def test_wolff_matches_exact_enumeration(orc, kind, dims, pbc, T):
    """Exchange only (the Wolff move ignores the field): long-run <E> and
    <|M|> of the oracle's Wolff steps equal the exact Boltzmann averages."""
    L = orc.lattice(kind, *dims, pbc=pbc)
    n = orc.num_sites(L)
    neigh = [[j for j in orc.neighbours(L, i) if j >= 0] for i in range(n)]
    M = orc.model(terms=("exchange",), temperature=T)
    spins = orc.ising_init(L, 2)
    orc.wolff_steps(L, M, 19, 0, 2000, spins)
    sums = orc.wolff_steps(L, M, 19, 2000, 60000, spins)
    bond, mag = sums[:, 0].astype(float), sums[:, 3].astype(float)
    e = -bond  # Exchange alone: total = -J sum over bonds
    ee, em, _, _ = ising_exact(neigh, n, 1.0, 0.0, 1.0, T, compound=False)
    me, se = binned(e)
    mm, sm = binned(np.abs(mag))
    assert abs(me - ee) < 4 * se, (me, ee, se)
    assert abs(mm - em) < 4 * sm, (mm, em, sm)


def test_wolff_clusters_do_not_depend_on_visit_order(orc):
    """The flipped set is the seed's component over active bonds: sums equal
    a recomputation from the state, and a second run is identical."""
    L = orc.lattice("sc", 6, 6, 6)
    M = orc.model(terms=("exchange",), temperature=4.0)
    a = orc.ising_init(L, 5)
    b = a.copy()
    sa = orc.wolff_steps(L, M, 7, 0, 50, a)
    sb = orc.wolff_steps(L, M, 7, 0, 50, b)
    assert np.array_equal(a, b) and np.array_equal(sa, sb)
    assert sa[-1, 3] == a.astype(np.int64).sum()
    s = a.astype(np.int64).reshape(6, 6, 6)
    assert sa[-1, 0] == sum(int((s * np.roll(s, 1, axis=ax)).sum()) for ax in range(3))


@pytest.mark.parametrize("dims,pbc,proposal,h", [((8, 6, 4), (1, 1, 1), "flip", 0.0),
                                                 ((6, 5, 3), (1, 0, 0), "random", 0.3),
                                                 ((10, 10, 1), (1, 1, 0), "flip", 0.1)])
def test_fast_sc_checkerboard_equals_the_oracle(orc, dims, pbc, proposal, h):
    """The CPU baseline's fast sc checkerboard makes the oracle's draws: the
    same states and sums, on 1 and on 4 threads."""
    L = orc.lattice("sc", *dims, pbc=pbc)
    M = orc.model(proposal=proposal, terms=("exchange", "zeeman"), temperature=3.0, field=h)
    ref = orc.ising_init(L, 8)
    want = orc.ising_sweeps(L, M, 8, 5, 7, ref)
    for threads in (1, 4):
        s = orc.ising_init(L, 8)
        got = orc.ising_sweeps_sc_fast(L, M, 8, 5, 7, s, threads)
        assert np.array_equal(s, ref) and np.array_equal(got, want), threads
