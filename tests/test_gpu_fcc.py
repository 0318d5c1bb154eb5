"""fcc chains (the four sublattices as the colours of a four-pass sweep) vs
the CPU checkerboard oracle: Ising bit-exact after every sweep, Heisenberg
fp32 bit-exact, over periodic and open boundaries, odd, degenerate and
segment-straddling sizes, random and flip proposals, and in batches mixed
with sc and bcc chains and run by the job server."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(vg, orc, dims, pbc, spin, proposal, terms, K=0.0, seed=3):
    L = vg.lattice(vg.FCC, *dims, pbc=tuple(bool(p) for p in pbc))
    kw = {"zeeman": "zeeman" in terms}
    if "exchange" in terms:
        kw["exchange"] = 1.0
    if "anisotropy" in terms:
        kw["anisotropy"] = K
    d = vg.chain(L, vg.ISING if spin == "ising" else vg.HEISENBERG,
                 vg.FLIP if proposal == "flip" else vg.RANDOM, vg.hamiltonian(**kw), seed)
    OL = orc.lattice("fcc", *dims, pbc=pbc)
    return d, OL


ISING = [
    # dims, pbc, T, h, hdir_z, proposal, terms
    ((16, 16, 16), (1, 1, 1), 9.8, 0.0, 1, "flip", ("exchange",)),
    ((5, 3, 7), (1, 1, 1), 8.0, 0.2, 1, "flip", ("exchange", "zeeman")),
    ((17, 4, 3), (0, 1, 0), 6.0, 0.0, 1, "flip", ("exchange",)),
    ((40, 2, 2), (1, 0, 1), 4.0, 0.1, -1, "flip", ("exchange", "zeeman")),
    ((1, 5, 5), (1, 1, 1), 9.0, 0.0, 1, "flip", ("exchange",)),
    ((33, 1, 1), (1, 1, 1), 5.0, 0.0, 1, "flip", ("exchange",)),
    ((20, 6, 4), (0, 0, 0), 7.0, 0.3, 1, "random", ("exchange", "zeeman")),
    ((32, 8, 8), (1, 1, 1), 2.0, 0.0, 1, "flip", ("exchange",)),
    ((48, 5, 3), (1, 1, 1), 30.0, 0.5, 1, "random", ("exchange", "zeeman")),
    ((16, 3, 2), (0, 1, 1), 9.0, 0.0, 1, "flip", ("exchange",)),
]


@pytest.mark.parametrize("dims,pbc,T,h,hz,proposal,terms", ISING)
def test_fcc_ising_bit_exact(vg, orc, ctx, dims, pbc, T, h, hz, proposal, terms):
    d, OL = _pair(vg, orc, dims, pbc, "ising", proposal, terms, seed=23)
    OM = orc.model(spin="ising", proposal=proposal, terms=terms, temperature=T, field=h,
                   field_dir=(0.0, 0.0, float(hz)))
    spins = orc.ising_init(OL, 23)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        assert np.array_equal(b.get_ising(0), spins)
        b.set_thermostat(0, T, h, (0.0, 0.0, float(hz)))
        for k in range(3):
            got = b.run_sums(1)[0]
            want = orc.ising_sweeps(OL, OM, 23, k, 1, spins)
            assert np.array_equal(got, want), k
            assert np.array_equal(b.get_ising(0), spins), k
        got = b.run_sums(9)[0]
        want = orc.ising_sweeps(OL, OM, 23, 3, 9, spins)
        assert np.array_equal(got, want)
        assert np.array_equal(b.get_ising(0), spins)


@pytest.mark.parametrize("dims,pbc", [((16, 16, 16), (1, 1, 1)), ((21, 6, 13), (1, 1, 0)),
                                      ((9, 5, 3), (0, 0, 1)), ((18, 7, 1), (1, 0, 0))])
def test_fcc_tiles_give_identical_trajectories(vg, orc, ctx, dims, pbc):
    """The fcc Ising pass marches K = 1, 2, 4 or 8 z planes per work item
    (the rows a pair needs again on the next plane stay in registers): every
    tile, including partial last chunks and open z ends, reproduces the
    oracle."""
    d, OL = _pair(vg, orc, dims, pbc, "ising", "flip", ("exchange", "zeeman"), seed=6)
    OM = orc.model(spin="ising", proposal="flip", terms=("exchange", "zeeman"),
                   temperature=8.0, field=0.1, field_dir=(0.0, 0.0, 1.0))
    spins = orc.ising_init(OL, 6)
    want = orc.ising_sweeps(OL, OM, 6, 0, 5, spins)
    for tile in (1, 2, 4, 8):
        with vg.Batch(ctx, [d]) as b:
            b.set_tile(tile)
            assert b.tile() == tile
            b.randomize()
            b.set_thermostat(0, 8.0, 0.1)
            assert np.array_equal(b.run_sums(5)[0], want), tile
            assert np.array_equal(b.get_ising(0), spins), tile


def test_fcc_tie_queue_with_lanes_that_stop_early(vg, orc, ctx):
    """fcc 32x6x13 cells, open z, tile 8: 12 lanes of each pass's wave march
    8 planes and 12 the last 5; tie records past index 12 belong to lanes
    that stopped early, and all are settled (20 sweeps equal the oracle's)."""
    d, OL = _pair(vg, orc, (32, 6, 13), (1, 1, 0), "ising", "flip", ("exchange",), seed=13)
    OM = orc.model(spin="ising", proposal="flip", terms=("exchange",), temperature=9.8)
    spins = orc.ising_init(OL, 13)
    want = orc.ising_sweeps(OL, OM, 13, 0, 20, spins)
    with vg.Batch(ctx, [d]) as b:
        b.set_tile(8)
        b.randomize()
        b.set_thermostat(0, 9.8)
        assert np.array_equal(b.run_sums(20)[0], want)
        assert np.array_equal(b.get_ising(0), spins)


def test_fcc_ising_set_state_and_sums(vg, orc, ctx):
    """A state set through the C ABI in the natural fcc site order (4 * cell
    + sublattice) comes back unchanged, and the per-sweep sums are the bond
    and magnetization sums of the downloaded state."""
    dims = (6, 5, 4)
    d, OL = _pair(vg, orc, dims, (1, 1, 1), "ising", "flip", ("exchange",), seed=4)
    n = orc.num_sites(OL)
    rng = np.random.default_rng(1)
    st = rng.choice(np.array([-1, 1], dtype=np.int8), n)
    neigh = [[j for j in orc.neighbours(OL, i) if j >= 0] for i in range(n)]
    with vg.Batch(ctx, [d]) as b:
        b.set_ising(0, st)
        assert np.array_equal(b.get_ising(0), st)
        b.set_thermostat(0, 7.0)
        sums = b.run_sums(2)[0]
        s = b.get_ising(0).astype(np.int64)
        assert sums[-1, 3] == s.sum()
        assert sums[-1, 0] == sum(int(s[i]) * int(s[j]) for i in range(n) for j in neigh[i]) // 2


HEIS = [
    # dims, pbc, T, h, proposal, terms, K
    ((8, 8, 8), (1, 1, 1), 2.0, 0.05, "random", ("exchange", "zeeman", "anisotropy"), 0.1),
    ((5, 3, 7), (1, 1, 1), 1.0, 0.0, "flip", ("exchange",), 0.0),
    ((9, 4, 3), (0, 1, 1), 1.4, 0.2, "random", ("exchange", "zeeman"), 0.0),
    ((1, 3, 3), (1, 1, 1), 0.7, 0.0, "random", ("exchange", "anisotropy"), 0.5),
    ((6, 1, 2), (1, 1, 0), 3.0, 0.1, "random", ("exchange", "zeeman", "anisotropy"), 0.1),
    ((13, 2, 2), (0, 0, 0), 0.5, 0.3, "random", ("exchange", "zeeman"), 0.0),
]


@pytest.mark.parametrize("dims,pbc,T,h,proposal,terms,K", HEIS)
def test_fcc_heisenberg_bit_exact(vg, orc, ctx, dims, pbc, T, h, proposal, terms, K):
    d, OL = _pair(vg, orc, dims, pbc, "heisenberg", proposal, terms, K, seed=9)
    OM = orc.model(spin="heisenberg", proposal=proposal, terms=terms, anisotropy=K,
                   temperature=T, field=h)
    xyz = orc.heis_init(OL, 9)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        assert np.array_equal(b.get_heisenberg(0), xyz.astype(np.float64))
        b.set_thermostat(0, T, h)
        got = b.run_sums(5)[0]
        want = orc.heis_sweeps(OL, OM, 9, 0, 5, xyz)
        assert np.array_equal(got, want)
        assert np.array_equal(b.get_heisenberg(0), xyz.astype(np.float64))


def test_mixed_lattice_batch(vg, orc, ctx):
    """sc, bcc and fcc, Ising and Heisenberg chains in one batch (six kernel
    groups, two and four colour passes) evolve exactly as each alone."""
    specs = [("sc", "ising", (16, 8, 8), 4.0, 1), ("fcc", "ising", (8, 8, 8), 9.0, 2),
             ("bcc", "heisenberg", (6, 6, 6), 2.0, 3), ("fcc", "heisenberg", (5, 6, 4), 1.5, 4),
             ("bcc", "ising", (12, 4, 4), 5.0, 5), ("fcc", "ising", (7, 3, 5), 7.0, 6)]
    kinds = {"sc": vg.SC, "bcc": vg.BCC, "fcc": vg.FCC}
    descs, refs = [], []
    for kind, spin, dims, T, seed in specs:
        L = vg.lattice(kinds[kind], *dims)
        H = vg.hamiltonian(exchange=1.0, zeeman=True)
        descs.append(vg.chain(L, vg.ISING if spin == "ising" else vg.HEISENBERG,
                              vg.FLIP if spin == "ising" else vg.RANDOM, H, seed))
        OL = orc.lattice(kind, *dims)
        OM = orc.model(spin=spin, proposal="flip" if spin == "ising" else "random",
                       terms=("exchange", "zeeman"), temperature=T, field=0.1)
        refs.append((spin, OL, OM, seed, T))
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        for i, (*_, T) in enumerate(refs):
            b.set_thermostat(i, T, 0.1)
        b.set_graph(True)
        got = b.run_sums(6)
        for i, (spin, OL, OM, seed, _) in enumerate(refs):
            if spin == "ising":
                st = orc.ising_init(OL, seed)
                want = orc.ising_sweeps(OL, OM, seed, 0, 6, st)
                assert np.array_equal(b.get_ising(i), st), i
            else:
                st = orc.heis_init(OL, seed)
                want = orc.heis_sweeps(OL, OM, seed, 0, 6, st)
                assert np.array_equal(b.get_heisenberg(i), st.astype(np.float64)), i
            assert np.array_equal(got[i], want), i


def test_fcc_observables_follow_reference_total_energy(vg, orc, ctx):
    """E and |M| per sweep equal the reference's total_energy and
    magnetization of the downloaded fcc state (Compound double counts)."""
    d, OL = _pair(vg, orc, (4, 3, 3), (1, 1, 0), "ising", "flip", ("exchange", "zeeman"), seed=5)
    OM = orc.model(terms=("exchange", "zeeman"), temperature=6.0, field=0.4)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        b.set_thermostat(0, 6.0, 0.4)
        e, m = b.run(4)
        s = orc.Sequential(OL, OM, 1)
        s.set_thermostat(6.0, 0.4)
        s.set_state(b.get_ising(0))
        assert e[0, -1] == pytest.approx(s.total_energy(), abs=1e-9)
        assert m[0, -1] == s.magnetization()


def test_fcc_jobs_served_alone_equal_in_mix(vg, ctx):
    """fcc jobs through the job server: each one's stage results equal a
    lone run of the same job, also when sc and bcc jobs share its blocks."""
    stages = vg.relax(steps=40, temperature=7.0) + vg.cool_down(8.0, 7.0, 0.5, 10, 30)
    jobs = [vg.chain(vg.lattice(vg.FCC, 6, 6, 6), vg.ISING, vg.FLIP,
                     vg.hamiltonian(exchange=1.0, zeeman=True), 31),
            vg.chain(vg.lattice(vg.SC, 8, 8, 8), vg.ISING, vg.FLIP,
                     vg.hamiltonian(exchange=1.0, zeeman=True), 32),
            vg.chain(vg.lattice(vg.FCC, 4, 5, 3), vg.HEISENBERG, vg.RANDOM,
                     vg.hamiltonian(exchange=1.0, zeeman=True), 33),
            vg.chain(vg.lattice(vg.BCC, 5, 5, 5), vg.HEISENBERG, vg.RANDOM,
                     vg.hamiltonian(exchange=1.0, zeeman=True), 34)]
    alone = []
    for d in jobs:
        with vg.Server(ctx, memory_bytes=64 << 20, max_jobs=1, quantum_sweeps=16) as srv:
            jid = srv.submit(d, stages)
            assert srv.wait(jid, 60.0).state == vg.JOB_DONE
            alone.append(srv.results(jid))
    with vg.Server(ctx, memory_bytes=64 << 20, max_jobs=4, quantum_sweeps=16) as srv:
        ids = [srv.submit(d, stages) for d in jobs]
        for k, jid in enumerate(ids):
            assert srv.wait(jid, 60.0).state == vg.JOB_DONE
            assert np.array_equal(srv.results(jid), alone[k]), k
