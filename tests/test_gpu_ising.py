"""Int8 Ising engine vs the CPU checkerboard oracle: identical spins after
every sweep and identical per-sweep sums (bit-exact), through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pbc(p):
    return tuple(bool(v) for v in p)


def make_pair(vg, orc, dims, pbc=(1, 1, 1), T=4.5115, h=0.0, hdir=(0.0, 0.0, 1.0),
              proposal="flip", terms=("exchange",), seed=7):
    L = vg.lattice(vg.SC, *dims, pbc=_pbc(pbc))
    kw = {}
    if "exchange" in terms:
        kw["exchange"] = 1.0
    H = vg.hamiltonian(zeeman="zeeman" in terms, **kw)
    d = vg.chain(L, vg.ISING, vg.FLIP if proposal == "flip" else vg.RANDOM, H, seed)
    OL = orc.lattice("sc", *dims, pbc=pbc)
    OM = orc.model(spin="ising", proposal=proposal, terms=terms, temperature=T, field=h,
                   field_dir=hdir)
    return d, OL, OM


def test_randomize_matches_oracle(vg, orc, ctx):
    for dims, pbc in [((64, 64, 64), (1, 1, 1)), ((10, 10, 1), (1, 1, 0)), ((34, 3, 5), (0, 0, 0))]:
        d, OL, _ = make_pair(vg, orc, dims, pbc, seed=42)
        with vg.Batch(ctx, [d]) as b:
            b.randomize()
            assert np.array_equal(b.get_ising(0), orc.ising_init(OL, 42)), dims


def test_set_get_round_trip(vg, orc, ctx):
    d, OL, _ = make_pair(vg, orc, (18, 6, 4), (1, 1, 1))
    rng = np.random.default_rng(1)
    s = rng.choice(np.array([-1, 1], dtype=np.int8), size=18 * 6 * 4)
    with vg.Batch(ctx, [d]) as b:
        b.set_ising(0, s)
        assert np.array_equal(b.get_ising(0), s)
        with pytest.raises(vg.VegasError):
            b.set_ising(0, np.zeros_like(s))


def test_64_cube_bit_exact_after_every_sweep(vg, orc, ctx):
    """sc 64^3, full PBC, J = 1, h = 0, T = 4.5115 (the 3D Ising critical
    point), seed 7: spins equal after each of 8 single-sweep calls, then
    after a 24-sweep call; every per-sweep sum equal."""
    d, OL, OM = make_pair(vg, orc, (64, 64, 64))
    spins = orc.ising_init(OL, 7)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        b.set_thermostat(0, 4.5115)
        for k in range(8):
            got = b.run_sums(1)[0]
            want = orc.ising_sweeps(OL, OM, 7, k, 1, spins)
            assert np.array_equal(got, want), k
            assert np.array_equal(b.get_ising(0), spins), k
        got = b.run_sums(24)[0]
        want = orc.ising_sweeps(OL, OM, 7, 8, 24, spins)
        assert np.array_equal(got, want)
        assert np.array_equal(b.get_ising(0), spins)
        assert b.sweep_count(0) == 32


CASES = [
    # dims, pbc, T, h, hdir, proposal, terms
    ((10, 10, 1), (1, 1, 0), 2.0, 0.0, (0, 0, 1), "flip", ("exchange", "zeeman")),
    ((2, 4, 6), (1, 1, 1), 3.0, 0.0, (0, 0, 1), "flip", ("exchange",)),
    ((40, 6, 4), (0, 1, 1), 3.5, 0.0, (0, 0, 1), "flip", ("exchange",)),
    ((10, 3, 5), (1, 0, 0), 2.5, 0.0, (0, 0, 1), "flip", ("exchange",)),
    ((34, 8, 8), (1, 1, 1), 4.0, 0.3, (0, 0, 1), "flip", ("exchange", "zeeman")),
    ((34, 8, 8), (1, 1, 1), 4.0, -0.3, (0, 0, -1), "flip", ("exchange", "zeeman")),
    ((32, 16, 8), (1, 1, 1), 1.0, 0.0, (0, 0, 1), "flip", ("exchange",)),
    ((32, 16, 8), (1, 1, 1), 50.0, 0.0, (0, 0, 1), "flip", ("exchange",)),
    ((16, 16, 16), (1, 1, 1), 4.0, 0.1, (0, 0, 1), "random", ("exchange", "zeeman")),
    ((36, 6, 2), (0, 0, 0), 3.0, 0.2, (0, 0, 1), "random", ("exchange", "zeeman")),
    ((12, 12, 12), (1, 1, 1), 1.3, 0.5, (0, 0, 1), "flip", ("zeeman",)),
]


@pytest.mark.parametrize("dims,pbc,T,h,hdir,proposal,terms", CASES)
def test_edge_cases_bit_exact(vg, orc, ctx, dims, pbc, T, h, hdir, proposal, terms):
    d, OL, OM = make_pair(vg, orc, dims, pbc, T, h, hdir, proposal, terms, seed=99)
    spins = orc.ising_init(OL, 99)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        b.set_thermostat(0, T, h, hdir)
        got = b.run_sums(6)[0]
        want = orc.ising_sweeps(OL, OM, 99, 0, 6, spins)
        assert np.array_equal(got, want)
        assert np.array_equal(b.get_ising(0), spins)


def test_mixed_batch_equals_each_chain_alone(vg, orc, ctx):
    """Chains of different sizes, temperatures, fields and seeds packed in one
    arena evolve exactly as each would alone (per-chain Philox keys and
    integer reductions)."""
    specs = [((16, 16, 16), 3.5, 0.05, 0), ((24, 24, 24), 4.5, -0.1, 1),
             ((8, 8, 8), 5.5, 0.0, 2), ((48, 16, 10), 4.0, 0.02, 3), ((10, 10, 1), 2.0, 0.0, 4)]
    descs, oracles = [], []
    for dims, T, h, seed in specs:
        pbc = (1, 1, 0) if dims[2] == 1 else (1, 1, 1)
        d, OL, OM = make_pair(vg, orc, dims, pbc, T, abs(h), (0, 0, 1 if h >= 0 else -1),
                              "flip", ("exchange", "zeeman"), seed)
        descs.append(d)
        oracles.append((OL, OM, seed, T, abs(h), (0, 0, 1 if h >= 0 else -1)))
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        for i, (_, _, _, T, h, hd) in enumerate(oracles):
            b.set_thermostat(i, T, h, hd)
        got = b.run_sums(10)
        for i, (OL, OM, seed, *_rest) in enumerate(oracles):
            spins = orc.ising_init(OL, seed)
            want = orc.ising_sweeps(OL, OM, seed, 0, 10, spins)
            assert np.array_equal(got[i], want), i
            assert np.array_equal(b.get_ising(i), spins), i


def test_graph_replay_equals_plain_launches(vg, orc, ctx):
    descs = [make_pair(vg, orc, dims, seed=s)[0]
             for dims, s in [((32, 32, 32), 1), ((16, 16, 16), 2), ((64, 8, 8), 3)]]
    results = []
    for graph in (False, True):
        with vg.Batch(ctx, descs) as b:
            b.set_graph(graph)
            b.randomize()
            for i in range(3):
                b.set_thermostat(i, 3.0 + i)
            sums = [b.run_sums(20) for _ in range(3)]  # the graph is replayed
            results.append((np.concatenate(sums, axis=1), [b.get_ising(i) for i in range(3)]))
    assert np.array_equal(results[0][0], results[1][0])
    for a, c in zip(results[0][1], results[1][1]):
        assert np.array_equal(a, c)


def test_prepare_runs_no_sweep(vg, orc, ctx):
    """vg_batch_prepare captures and uploads the graphs of a run without
    running a sweep: counters and spins stay, and the prepared run equals the
    oracle (a full block of 256 sweeps and a remainder of 44)."""
    d, OL, OM = make_pair(vg, orc, (32, 8, 8), seed=2)
    spins = orc.ising_init(OL, 2)
    with vg.Batch(ctx, [d]) as b:
        b.set_graph(True)
        b.randomize()
        b.set_thermostat(0, 4.5115)
        before = b.get_ising(0)
        b.prepare(300)
        assert b.sweep_count(0) == 0 and np.array_equal(b.get_ising(0), before)
        want = orc.ising_sweeps(OL, OM, 2, 0, 300, spins)
        assert np.array_equal(b.run_sums(300)[0], want)
        assert np.array_equal(b.get_ising(0), spins)


def test_observables_follow_reference_total_energy(vg, orc, ctx):
    """E and |M| returned per sweep equal the reference's total_energy and
    magnetization of the downloaded state (computed by the oracle's
    site-by-site sum over the CSR matrix): Compound double counts exchange,
    Exchange alone halves it (Q4)."""
    for terms in [("exchange", "zeeman"), ("exchange",), ("zeeman",)]:
        d, OL, OM = make_pair(vg, orc, (10, 10, 1), (1, 1, 0), 2.2, 0.4, (0, 0, 1), "flip",
                              terms, seed=5)
        with vg.Batch(ctx, [d]) as b:
            b.randomize()
            b.set_thermostat(0, 2.2, 0.4)
            e, m = b.run(5)
            s = orc.Sequential(OL, OM, 1)
            s.set_thermostat(2.2, 0.4)
            s.set_state(b.get_ising(0))
            assert e[0, -1] == pytest.approx(s.total_energy(), abs=1e-9)
            assert m[0, -1] == s.magnetization()


def test_256_cube_sums_match_downloaded_state(vg, ctx):
    """At the full 256^3 size: the reported magnetization and bond sum equal
    the ones recomputed from the downloaded spins."""
    L = vg.lattice(vg.SC, 256, 256, 256)
    d = vg.chain(L, vg.ISING, vg.FLIP, vg.hamiltonian(exchange=1.0), 7)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        b.set_thermostat(0, 4.5115)
        sums = b.run_sums(3)[0]
        s = b.get_ising(0).astype(np.int64).reshape(256, 256, 256)  # [z][y][x]
        assert sums[-1, 3] == s.sum()
        bond = sum(int((s * np.roll(s, 1, axis=a)).sum()) for a in range(3))
        assert sums[-1, 0] == bond
        assert sums[-1, 4] == 256 ** 3


def test_errors_mirror_reference_kinds(vg, ctx):
    L = vg.lattice(vg.SC, 5, 4, 4)  # odd periodic x: not bipartite
    with pytest.raises(vg.VegasError) as e:
        vg.Batch(ctx, [vg.chain(L)])
    assert e.value.kind == "UnsupportedLattice"
    L = vg.lattice(vg.SC, 4, 4, 4)
    bad = vg.chain(L, hamiltonian=vg.Hamiltonian(0, 0, 0.0, 0.0, (0.0, 0.0, 1.0), 0.0))
    with pytest.raises(vg.VegasError):
        vg.Batch(ctx, [bad])


@pytest.mark.parametrize("dims,pbc", [((64, 64, 64), (1, 1, 1)), ((34, 6, 13), (1, 1, 0)),
                                      ((10, 10, 1), (1, 1, 0))])
def test_tiles_give_identical_trajectories(vg, orc, ctx, dims, pbc):
    """The z tile (planes per work item) changes only the work layout: every
    tile reproduces the oracle bit for bit."""
    d, OL, OM = make_pair(vg, orc, dims, pbc, 3.0, 0.1, (0, 0, 1), "flip",
                          ("exchange", "zeeman"), seed=5)
    spins = orc.ising_init(OL, 5)
    want = orc.ising_sweeps(OL, OM, 5, 0, 4, spins)
    for tile in (1, 2, 4, 8):
        with vg.Batch(ctx, [d]) as b:
            b.set_tile(tile)
            assert b.tile() == tile
            b.randomize()
            b.set_thermostat(0, 3.0, 0.1)
            assert np.array_equal(b.run_sums(4)[0], want), tile
            assert np.array_equal(b.get_ising(0), spins), tile


def test_tie_queue_with_lanes_that_stop_early(vg, orc, ctx):
    """sc 64x6x13, open z, tile 8: the one wave of each pass holds 12 lanes
    marching 8 planes and 12 marching the last 5, and queues about ten tie
    records per pass -- records past index 12 belong to lanes that stopped
    early.  Every record is settled: 30 sweeps equal the oracle's."""
    d, OL, OM = make_pair(vg, orc, (64, 6, 13), (1, 1, 0), 4.5115, 0.0, (0, 0, 1), "flip",
                          ("exchange",), seed=13)
    spins = orc.ising_init(OL, 13)
    want = orc.ising_sweeps(OL, OM, 13, 0, 30, spins)
    with vg.Batch(ctx, [d]) as b:
        b.set_tile(8)
        b.randomize()
        b.set_thermostat(0, 4.5115)
        assert np.array_equal(b.run_sums(30)[0], want)
        assert np.array_equal(b.get_ising(0), spins)


def test_autotune_leaves_trajectory_unchanged(vg, orc, ctx):
    descs = [make_pair(vg, orc, (32, 32, 32), seed=s)[0] for s in (1, 2)]
    runs = []
    for tune in (False, True):
        with vg.Batch(ctx, descs) as b:
            b.randomize()
            for i in range(2):
                b.set_thermostat(i, 4.0)
            b.run(3)
            if tune:
                assert b.autotune(4) in (1, 2, 4, 8)
            runs.append((b.run_sums(5), [b.get_ising(i) for i in range(2)],
                         [b.sweep_count(i) for i in range(2)]))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert all(np.array_equal(a, c) for a, c in zip(runs[0][1], runs[1][1]))
    assert runs[0][2] == runs[1][2] == [8, 8]


def test_sweep_counter_limit(vg, ctx):
    """Philox takes 56 bits of the sweep counter: setting or running past
    2^56 is refused, up to it works and continues the counter."""
    d = vg.chain(vg.lattice(vg.SC, 8, 8, 8), vg.ISING, vg.FLIP, vg.hamiltonian(exchange=1.0), 3)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        with pytest.raises(vg.VegasError):
            b.set_sweep_count(0, 1 << 56)
        b.set_sweep_count(0, (1 << 56) - 4)
        b.run(4)
        assert b.sweep_count(0) == 1 << 56
        with pytest.raises(vg.VegasError):
            b.run(1)
