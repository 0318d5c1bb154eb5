"""Wolff cluster steps (reference WolffIntegrator, src/integrator.rs:146-230)
on the GPU vs the CPU oracle: states and per-step sums bit-exact on sc, bcc
and fcc lattices with periodic and open axes, over several sweep blocks and
in batches mixed with Metropolis chains; observables against the
reference's total_energy; 3-sigma agreement with the sequential Metropolis
restatement; the refused combinations."""
import numpy as np
import pytest

from _stats import binned, z_score

pytestmark = pytest.mark.gpu

KINDS = {"sc": 0, "bcc": 1, "fcc": 2}


def _chain(vg, kind, dims, pbc, seed, zeeman=False):
    L = vg.lattice([vg.SC, vg.BCC, vg.FCC][KINDS[kind]], *dims,
                   pbc=tuple(bool(p) for p in pbc))
    return vg.chain(L, vg.ISING, vg.WOLFF, vg.hamiltonian(exchange=1.0, zeeman=zeeman), seed)


CASES = [
    # kind, dims, pbc, T
    ("sc", (16, 16, 16), (1, 1, 1), 4.5115),
    ("sc", (10, 10, 1), (1, 1, 0), 2.0),
    ("sc", (8, 6, 4), (0, 1, 1), 6.0),
    ("sc", (24, 24, 24), (1, 1, 1), 3.0),
    ("bcc", (7, 5, 6), (1, 1, 1), 6.3),
    ("bcc", (6, 6, 3), (1, 0, 1), 4.0),
    ("fcc", (5, 4, 6), (1, 1, 1), 9.8),
    ("fcc", (4, 4, 4), (0, 0, 1), 12.0),
    ("sc", (2, 2, 2), (1, 1, 1), 1.0),
]


@pytest.mark.parametrize("kind,dims,pbc,T", CASES)
def test_wolff_bit_exact(vg, orc, ctx, kind, dims, pbc, T):
    d = _chain(vg, kind, dims, pbc, 41)
    OL = orc.lattice(kind, *dims, pbc=pbc)
    OM = orc.model(terms=("exchange",), temperature=T)
    spins = orc.ising_init(OL, 41)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        assert np.array_equal(b.get_ising(0), spins)
        b.set_thermostat(0, T)
        step = 0
        for n in (1, 7, 300):  # the last spans two sweep blocks
            got = b.run_sums(n)[0]
            want = orc.wolff_steps(OL, OM, 41, step, n, spins)
            step += n
            assert np.array_equal(got, want), n
            assert np.array_equal(b.get_ising(0), spins), n
        assert b.sweep_count(0) == step


def test_wolff_set_state_and_batches_with_metropolis_chains(vg, orc, ctx):
    """Wolff chains next to checkerboard chains in one batch evolve exactly
    as alone; a state set through the C ABI is the one the steps start from."""
    specs = [("wolff", "sc", (12, 12, 12), 4.4, 1), ("flip", "sc", (16, 8, 8), 4.0, 2),
             ("wolff", "fcc", (4, 5, 3), 9.0, 3), ("flip", "bcc", (6, 6, 6), 6.0, 4),
             ("wolff", "bcc", (5, 5, 5), 5.5, 5)]
    descs, refs = [], []
    rng = np.random.default_rng(3)
    for alg, kind, dims, T, seed in specs:
        L = vg.lattice([vg.SC, vg.BCC, vg.FCC][KINDS[kind]], *dims)
        descs.append(vg.chain(L, vg.ISING, vg.WOLFF if alg == "wolff" else vg.FLIP,
                              vg.hamiltonian(exchange=1.0, zeeman=True), seed))
        OL = orc.lattice(kind, *dims)
        st = rng.choice(np.array([-1, 1], dtype=np.int8), orc.num_sites(OL))
        refs.append((alg, OL, orc.model(terms=("exchange", "zeeman"), temperature=T, field=0.2),
                     seed, T, st))
    with vg.Batch(ctx, descs) as b:
        for i, (*_, T, st) in enumerate(refs):
            b.set_ising(i, st)
            assert np.array_equal(b.get_ising(i), st)
            b.set_thermostat(i, T, 0.2)
        b.set_graph(True)
        got = b.run_sums(40)
        for i, (alg, OL, OM, seed, _, st) in enumerate(refs):
            s = st.copy()
            if alg == "wolff":
                want = orc.wolff_steps(OL, OM, seed, 0, 40, s)
            else:
                want = orc.ising_sweeps(OL, OM, seed, 0, 40, s)
            assert np.array_equal(got[i], want), i
            assert np.array_equal(b.get_ising(i), s), i


def test_wolff_lds_and_global_paths(vg, orc, ctx):
    """Chains of at most 2^19 sites keep spins and visit marks in LDS
    (wolff_run_lds), larger ones run from global memory (wolff_run); both
    kinds in one batch, and the small one alone, match the oracle's steps
    bit for bit over two sweep blocks, at T_c (large clusters) and at a low
    temperature (the cluster covers most of the lattice)."""
    specs = [("sc", (82, 82, 80), 4.5115, 21),   # 537920 sites: global memory
             ("sc", (80, 80, 80), 3.0, 22),      # 512000 sites: LDS, clusters of most sites
             ("bcc", (64, 64, 65), 3.0, 23),     # 532480 sites: global memory
             ("fcc", (8, 8, 8), 2.0, 24),        # 2048 sites at low T: LDS
             ("sc", (46, 42, 34), 4.5115, 25)]   # LDS, a site count not a multiple of 32
    descs, refs = [], []
    for kind, dims, T, seed in specs:
        descs.append(_chain(vg, kind, dims, (1, 1, 1), seed))
        OL = orc.lattice(kind, *dims)
        refs.append((OL, orc.model(terms=("exchange",), temperature=T), seed, T,
                     orc.ising_init(OL, seed)))
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        for i, (_, _, _, T, _) in enumerate(refs):
            b.set_thermostat(i, T)
        step = 0
        for n in (3, 40):
            got = b.run_sums(n)
            for i, (OL, OM, seed, _, st) in enumerate(refs):
                want = orc.wolff_steps(OL, OM, seed, step, n, st)
                assert np.array_equal(got[i], want), (i, n)
                assert np.array_equal(b.get_ising(i), st), (i, n)
            step += n
    # the LDS chain alone continues as in the mix
    OL, OM, seed, T, _ = refs[1]
    st = orc.ising_init(OL, seed)
    with vg.Batch(ctx, [descs[1]]) as b:
        b.randomize()
        b.set_thermostat(0, T)
        got = b.run_sums(43)[0]
        assert np.array_equal(got, orc.wolff_steps(OL, OM, seed, 0, 43, st))
        assert np.array_equal(b.get_ising(0), st)


def test_wolff_observables_follow_reference_total_energy(vg, orc, ctx):
    """E and |M| per step are the reference's Compound total energy and
    magnetization of the state (the field enters E, not the moves)."""
    d = _chain(vg, "sc", (8, 8, 8), (1, 1, 1), 9, zeeman=True)
    OL = orc.lattice("sc", 8, 8, 8)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        b.set_thermostat(0, 4.0, 0.3)
        e, m = b.run(5)
        s = orc.Sequential(OL, orc.model(terms=("exchange", "zeeman"), temperature=4.0, field=0.3), 1)
        s.set_thermostat(4.0, 0.3)
        s.set_state(b.get_ising(0))
        assert e[0, -1] == pytest.approx(s.total_energy(), abs=1e-9)
        assert m[0, -1] == s.magnetization()


def test_wolff_agrees_with_sequential_metropolis(vg, orc, ctx):
    """sc 8^3 at T = 4.5 (near T_c), Exchange only: <E>/N and <|M|>/N of
    Wolff chains (many steps, independent seeds) agree within 3 sigma with
    the sequential random-site Metropolis restatement."""
    T, n, chains, steps = 4.5, 512, 64, 3000
    with vg.Batch(ctx, [_chain(vg, "sc", (8, 8, 8), (1, 1, 1), s) for s in range(chains)]) as b:
        b.randomize()
        for i in range(chains):
            b.set_thermostat(i, T)
        b.run(500, observe=False)
        e, m = b.run(steps)
    ge = (e.mean(-1) / n).mean(), (e.mean(-1) / n).std(ddof=1) / np.sqrt(chains)
    gm = (m.mean(-1) / n).mean(), (m.mean(-1) / n).std(ddof=1) / np.sqrt(chains)
    OL = orc.lattice("sc", 8, 8, 8)
    s = orc.Sequential(OL, orc.model(terms=("exchange",), temperature=T), 77)
    s.set_thermostat(T, 0.0)
    s.set_state(orc.ising_init(OL, 77))
    s.sweeps(2000, observe=False)
    se, sm = s.sweeps(60000)
    s.close()
    for (a, ea), (bb, eb), lab in ((ge, binned(se / n), "E/N"), (gm, binned(sm / n), "|M|/N")):
        z = z_score(a, ea, bb, eb)
        print(f"wolff {lab}: {a:.5f} +- {ea:.5f}  sequential {bb:.5f} +- {eb:.5f}  z={z:.2f}")
        assert z < 3.0, (lab, a, ea, bb, eb)


def test_wolff_refused_combinations(vg, ctx):
    L = vg.lattice(vg.SC, 4, 4, 4)
    for spin in (vg.HEISENBERG, vg.ISING_BITS):
        with pytest.raises(vg.VegasError) as e:
            vg.Batch(ctx, [vg.chain(L, spin, vg.WOLFF, vg.hamiltonian(exchange=1.0), 1)])
        assert e.value.kind == "NotImplemented"
    # 2^31 sites: past the 31-bit site ids of the Wolff kernels (refused
    # before anything is allocated)
    big = vg.lattice(vg.SC, 2048, 1024, 1024)
    with pytest.raises(vg.VegasError) as e:
        vg.Batch(ctx, [vg.chain(big, vg.ISING, vg.WOLFF, vg.hamiltonian(exchange=1.0), 1)])
    assert e.value.kind == "InvalidArgument"
    with vg.Server(ctx, memory_bytes=1 << 20, max_jobs=2, quantum_sweeps=10) as srv:
        with pytest.raises(vg.VegasError) as e:
            srv.submit(vg.chain(L, vg.ISING, vg.WOLFF, vg.hamiltonian(exchange=1.0), 1),
                       vg.relax(steps=10, temperature=3.0))
        assert e.value.kind == "NotImplemented"
