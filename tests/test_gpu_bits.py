"""Bit-packed Ising chains (1 bit per site) on sc, bcc and fcc lattices: the same
trajectories as the int8 chains and the CPU oracle, bit for bit, over sizes
whose rows end inside a 16-position segment, open and periodic axes, both
proposals and fields."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # dims, pbc, T, h, proposal, terms
    ((64, 64, 64), (1, 1, 1), 4.5115, 0.0, "flip", ("exchange",)),
    ((34, 6, 4), (1, 1, 1), 3.0, 0.2, "flip", ("exchange", "zeeman")),
    ((10, 10, 1), (1, 1, 0), 2.2, 0.0, "flip", ("exchange", "zeeman")),
    ((2, 4, 6), (1, 1, 1), 3.0, 0.0, "flip", ("exchange",)),
    ((40, 6, 4), (0, 1, 1), 3.5, 0.1, "random", ("exchange", "zeeman")),
    ((66, 3, 5), (0, 0, 0), 2.5, 0.0, "flip", ("exchange",)),
    ((48, 16, 10), (1, 1, 1), 1.0, 0.0, "flip", ("exchange",)),
    ((16, 16, 16), (1, 1, 1), 4.0, -0.3, "random", ("exchange", "zeeman")),
]


def _desc(vg, dims, pbc, spin, proposal, terms, seed, kind=None):
    L = vg.lattice(vg.SC if kind is None else kind, *dims, pbc=tuple(bool(p) for p in pbc))
    kw = {"zeeman": "zeeman" in terms}
    if "exchange" in terms:
        kw["exchange"] = 1.0
    return vg.chain(L, spin, vg.FLIP if proposal == "flip" else vg.RANDOM,
                    vg.hamiltonian(**kw), seed)


@pytest.mark.parametrize("dims,pbc,T,h,proposal,terms", CASES)
def test_bits_match_oracle_and_int8(vg, orc, ctx, dims, pbc, T, h, proposal, terms):
    hd = (0.0, 0.0, 1.0 if h >= 0 else -1.0)
    OL = orc.lattice("sc", *dims, pbc=pbc)
    OM = orc.model(spin="ising", proposal=proposal, terms=terms, temperature=T, field=abs(h),
                   field_dir=hd)
    spins = orc.ising_init(OL, 31)
    descs = [_desc(vg, dims, pbc, vg.ISING_BITS, proposal, terms, 31),
             _desc(vg, dims, pbc, vg.ISING, proposal, terms, 31)]
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        assert np.array_equal(b.get_ising(0), spins)
        for i in range(2):
            b.set_thermostat(i, T, abs(h), hd)
        for k in range(2):
            got = b.run_sums(1)
            want = orc.ising_sweeps(OL, OM, 31, k, 1, spins)
            assert np.array_equal(got[0], want) and np.array_equal(got[1], want), k
            assert np.array_equal(b.get_ising(0), spins), k
        got = b.run_sums(10)
        want = orc.ising_sweeps(OL, OM, 31, 2, 10, spins)
        assert np.array_equal(got[0], want) and np.array_equal(got[1], want)
        assert np.array_equal(b.get_ising(0), spins)
        assert np.array_equal(b.get_ising(1), spins)


BCC_CASES = [
    # dims (cells), pbc, T, h, proposal, terms
    ((16, 16, 16), (1, 1, 1), 6.35, 0.0, "flip", ("exchange",)),
    ((17, 4, 3), (0, 1, 0), 4.0, 0.2, "flip", ("exchange", "zeeman")),
    ((33, 5, 7), (1, 1, 1), 5.0, 0.0, "random", ("exchange", "zeeman")),
    ((40, 2, 2), (1, 0, 1), 3.0, -0.1, "flip", ("exchange", "zeeman")),
    ((5, 3, 13), (0, 0, 0), 30.0, 0.5, "random", ("exchange", "zeeman")),
]


FCC_CASES = [
    # dims (cells), pbc, T, h, proposal, terms
    ((16, 16, 16), (1, 1, 1), 9.8, 0.0, "flip", ("exchange",)),
    ((17, 4, 3), (0, 1, 0), 6.0, 0.2, "flip", ("exchange", "zeeman")),
    ((33, 5, 7), (1, 1, 1), 8.0, 0.0, "random", ("exchange", "zeeman")),
    ((40, 2, 2), (1, 0, 1), 4.0, -0.1, "flip", ("exchange", "zeeman")),
    ((5, 3, 13), (0, 0, 0), 30.0, 0.5, "random", ("exchange", "zeeman")),
    ((1, 5, 5), (1, 1, 1), 9.0, 0.0, "flip", ("exchange",)),
    ((32, 8, 8), (1, 1, 1), 2.0, 0.0, "flip", ("exchange",)),
]


def _bits_vs_int8(vg, orc, ctx, lattice, kind, dims, pbc, T, h, proposal, terms):
    """The bit-packed and int8 chains of one lattice against the oracle, over
    every tile."""
    hd = (0.0, 0.0, 1.0 if h >= 0 else -1.0)
    OL = orc.lattice(lattice, *dims, pbc=pbc)
    OM = orc.model(spin="ising", proposal=proposal, terms=terms, temperature=T, field=abs(h),
                   field_dir=hd)
    spins = orc.ising_init(OL, 17)
    descs = [_desc(vg, dims, pbc, vg.ISING_BITS, proposal, terms, 17, kind),
             _desc(vg, dims, pbc, vg.ISING, proposal, terms, 17, kind)]
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        assert np.array_equal(b.get_ising(0), spins)
        for i in range(2):
            b.set_thermostat(i, T, abs(h), hd)
        for tile, n in [(4, 2), (1, 3), (2, 2), (8, 4)]:
            b.set_tile(tile)
            got = b.run_sums(n)
            want = orc.ising_sweeps(OL, OM, 17, b.sweep_count(0) - n, n, spins)
            assert np.array_equal(got[0], want) and np.array_equal(got[1], want), tile
            assert np.array_equal(b.get_ising(0), spins), tile
            assert np.array_equal(b.get_ising(1), spins), tile


@pytest.mark.parametrize("dims,pbc,T,h,proposal,terms", BCC_CASES)
def test_bcc_bits_match_oracle_and_int8(vg, orc, ctx, dims, pbc, T, h, proposal, terms):
    _bits_vs_int8(vg, orc, ctx, "bcc", vg.BCC, dims, pbc, T, h, proposal, terms)


@pytest.mark.parametrize("dims,pbc,T,h,proposal,terms", FCC_CASES)
def test_fcc_bits_match_oracle_and_int8(vg, orc, ctx, dims, pbc, T, h, proposal, terms):
    _bits_vs_int8(vg, orc, ctx, "fcc", vg.FCC, dims, pbc, T, h, proposal, terms)


@pytest.mark.parametrize("lattice,T,n", [("bcc", 6.35, 30), ("fcc", 9.8, 20)])
def test_bits_tie_queue_with_lanes_that_stop_early(vg, orc, ctx, lattice, T, n):
    """32x6x13 cells, open z, tile 8: 12 lanes of each pass's wave march 8
    planes and 12 the last 5; the tie records of the bit-packed passes
    (settled by rewriting the owner's 16-bit word) are all applied."""
    d = _desc(vg, (32, 6, 13), (1, 1, 0), vg.ISING_BITS, "flip", ("exchange",), 13,
              getattr(vg, lattice.upper()))
    OL = orc.lattice(lattice, 32, 6, 13, pbc=(1, 1, 0))
    OM = orc.model(spin="ising", proposal="flip", terms=("exchange",), temperature=T)
    spins = orc.ising_init(OL, 13)
    want = orc.ising_sweeps(OL, OM, 13, 0, n, spins)
    with vg.Batch(ctx, [d]) as b:
        b.set_tile(8)
        b.randomize()
        b.set_thermostat(0, T)
        assert np.array_equal(b.run_sums(n)[0], want)
        assert np.array_equal(b.get_ising(0), spins)


@pytest.mark.parametrize("kind,ncol", [("BCC", 2), ("FCC", 4)])
def test_bits_set_get_natural_order(vg, ctx, kind, ncol):
    """A bcc or fcc state set in the natural site order (ncol * cell +
    sublattice) comes back unchanged from the bit planes."""
    d = _desc(vg, (21, 6, 5), (1, 1, 0), vg.ISING_BITS, "flip", ("exchange",), 5,
              getattr(vg, kind))
    rng = np.random.default_rng(8)
    s = rng.choice(np.array([-1, 1], dtype=np.int8), size=ncol * 21 * 6 * 5)
    with vg.Batch(ctx, [d]) as b:
        b.set_ising(0, s)
        assert np.array_equal(b.get_ising(0), s)


def test_bits_set_get_tiles_graph(vg, ctx):
    d = _desc(vg, (34, 6, 13), (1, 1, 0), vg.ISING_BITS, "flip", ("exchange", "zeeman"), 5)
    rng = np.random.default_rng(3)
    s = rng.choice(np.array([-1, 1], dtype=np.int8), size=34 * 6 * 13)
    runs = []
    for tile, graph in [(1, False), (2, True), (4, False), (8, True)]:
        with vg.Batch(ctx, [d]) as b:
            b.set_ising(0, s)
            assert np.array_equal(b.get_ising(0), s)
            b.set_tile(tile)
            b.set_graph(graph)
            b.set_thermostat(0, 3.0, 0.1)
            runs.append((b.run_sums(7), b.get_ising(0)))
    for r in runs[1:]:
        assert np.array_equal(r[0], runs[0][0]) and np.array_equal(r[1], runs[0][1])


def test_autotune_restores_every_group(vg, orc, ctx):
    """Tuning trials run every chain of the batch; all kernel groups' spins
    and counters come back, so trajectories are unchanged."""
    H = vg.hamiltonian(exchange=1.0, zeeman=True)
    Hh = vg.hamiltonian(exchange=1.0, zeeman=True, anisotropy=0.1)
    descs = [vg.chain(vg.lattice(vg.SC, 32, 32, 32), vg.ISING, vg.FLIP, H, 1),
             vg.chain(vg.lattice(vg.SC, 32, 32, 32), vg.ISING_BITS, vg.FLIP, H, 2),
             vg.chain(vg.lattice(vg.SC, 8, 8, 8), vg.HEISENBERG, vg.RANDOM, Hh, 3),
             vg.chain(vg.lattice(vg.BCC, 6, 6, 6), vg.ISING, vg.FLIP, H, 4),
             vg.chain(vg.lattice(vg.BCC, 18, 6, 6), vg.ISING_BITS, vg.FLIP, H, 5),
             vg.chain(vg.lattice(vg.FCC, 18, 6, 6), vg.ISING_BITS, vg.FLIP, H, 6)]
    runs = []
    for tune in (False, True):
        with vg.Batch(ctx, descs) as b:
            b.randomize()
            for i in range(6):
                b.set_thermostat(i, (3.0, 3.0, 1.2, 3.0, 3.0, 7.0)[i], 0.05)
            b.run(3)
            if tune:
                assert b.autotune(3) in (1, 2, 4, 8)
            runs.append((b.run_sums(5), b.get_ising(0), b.get_ising(1), b.get_heisenberg(2),
                         b.get_ising(3), b.get_ising(4), b.get_ising(5),
                         [b.sweep_count(i) for i in range(6)]))
    for a, c in zip(runs[0], runs[1]):
        assert np.array_equal(np.asarray(a), np.asarray(c))
    assert runs[1][-1] == [8] * 6


def test_bits_memory_is_an_eighth(vg, ctx):
    """sc 256^3: the bit-packed chain runs the same sweeps as the int8 one
    (identical sums) from planes one eighth the size."""
    L = vg.lattice(vg.SC, 256, 256, 256)
    H = vg.hamiltonian(exchange=1.0)
    with vg.Batch(ctx, [vg.chain(L, vg.ISING_BITS, vg.FLIP, H, 7),
                        vg.chain(L, vg.ISING, vg.FLIP, H, 7)]) as b:
        b.randomize()
        for i in range(2):
            b.set_thermostat(i, 4.5115)
        s = b.run_sums(3)
        assert np.array_equal(s[0], s[1])
        assert np.array_equal(b.get_ising(0), b.get_ising(1))
