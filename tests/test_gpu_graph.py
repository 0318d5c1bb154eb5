"""Chains whose exchange is an arbitrary sparse matrix (Exchange::new(CsMat),
src/energy.rs:170-172) on the GPU, against the oracle's restatement.

- bit-exact spins and per-sweep sums after every sweep: random graphs with
  mixed-sign couplings, duplicate, diagonal and one-sided entries, a field,
  flip and random proposals, Ising and Heisenberg, in one batch with lattice
  chains;
- the colouring the engine reports equals the oracle's;
- the matrix of an sc lattice gives the lattice chain's equilibrium (3 sigma
  against the sequential restatement over the same matrix);
- E from the sums equals the reference's total_energy of the downloaded
  state, and errors (bad CSR, unsupported kinds) are reported."""
import numpy as np
import pytest

from _stats import binned, z_score

pytestmark = pytest.mark.gpu


def random_csr(rng, n, p, symmetric=True, dup=False, diag=False, scale=1.0):
    rows = [[] for _ in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < p:
                v = float(rng.normal() * scale)
                u = v if symmetric else float(rng.normal() * scale)
                rows[i].append((j, v))
                if symmetric or rng.random() < 0.7:   # some one-sided entries
                    rows[j].append((i, u))
    row_ptr, col, val = [0], [], []
    for i in range(n):
        ent = list(rows[i])
        rng.shuffle(ent)                              # any column order
        if diag:
            ent.append((i, 0.25))
        for j, v in ent:
            if dup:
                col += [j, j]
                val += [0.5 * v, 0.5 * v]
            else:
                col.append(j)
                val.append(v)
        row_ptr.append(len(col))
    return (np.array(row_ptr, dtype=np.uint64), np.array(col, dtype=np.uint64),
            np.array(val, dtype=np.float64))


def lattice_csr(orc, kind, dims, pbc):
    L = orc.lattice(kind, *dims, pbc=pbc)
    n = orc.num_sites(L)
    row_ptr, col = [0], []
    for i in range(n):
        col += sorted(j for j in orc.neighbours(L, i) if j >= 0)
        row_ptr.append(len(col))
    return (np.array(row_ptr, dtype=np.uint64), np.array(col, dtype=np.uint64),
            np.ones(len(col)))


GRAPHS = [
    # n, p, symmetric, dup, diag, scale
    (37, 0.12, True, False, False, 1.0),
    (60, 0.08, False, True, True, 0.7),
    (300, 0.02, True, True, False, 2.0),
    (1000, 0.004, False, False, True, 1.0),
]


@pytest.mark.parametrize("spin", ["ising", "heisenberg"])
@pytest.mark.parametrize("proposal", ["flip", "random"])
def test_graph_chains_bit_exact(vg, orc, ctx, spin, proposal):
    rng = np.random.default_rng(7 if spin == "ising" else 8)
    mats = [random_csr(rng, *g) for g in GRAPHS]
    terms = ("exchange", "zeeman") if spin == "ising" else ("exchange", "zeeman", "anisotropy")
    H = vg.hamiltonian(exchange=1.0, zeeman=True,
                       anisotropy=0.2 if spin == "heisenberg" else None)
    sk = vg.ISING if spin == "ising" else vg.HEISENBERG
    pk = vg.FLIP if proposal == "flip" else vg.RANDOM
    descs = [vg.chain(vg.graph_lattice(), sk, pk, H, 100 + k) for k in range(len(mats))]
    # a lattice chain in the same batch
    descs.append(vg.chain(vg.lattice(vg.SC, 8, 6, 4), sk, pk, H, 99))
    exch = list(mats) + [None]
    T, h = (1.7, 0.3) if spin == "ising" else (0.9, 0.2)
    with vg.Batch(ctx, descs, exchange=exch) as b:
        b.randomize()
        for i in range(len(descs)):
            b.set_thermostat(i, T, h)
        graphs = [orc.Graph(*m) for m in mats]
        for k, G in enumerate(graphs):
            assert b.colours(k) == G.colours()
            sc = b.sums_scales(k)
            assert sc[0] == 2.0 ** -27
        OM = orc.model(spin=spin, proposal=proposal, terms=terms, anisotropy=0.2,
                       temperature=T, field=h)
        states = [G.ising_init(100 + k) if spin == "ising" else G.heis_init(100 + k)
                  for k, G in enumerate(graphs)]
        for k in range(len(graphs)):
            got = b.get_ising(k) if spin == "ising" else b.get_heisenberg(k)
            assert np.array_equal(got, states[k].astype(got.dtype)), k
        for sweep in range(5):
            print(f"graph {spin} {proposal} sweep {sweep}", flush=True)
            got = b.run_sums(1)
            for k, G in enumerate(graphs):
                run = G.ising_sweeps if spin == "ising" else G.heis_sweeps
                want = run(OM, 100 + k, sweep, 1, states[k])
                assert np.array_equal(got[k], want), (k, sweep, got[k], want)
                st = b.get_ising(k) if spin == "ising" else b.get_heisenberg(k)
                assert np.array_equal(st, states[k].astype(st.dtype)), (k, sweep)
        got = b.run_sums(40)
        for k, G in enumerate(graphs):
            run = G.ising_sweeps if spin == "ising" else G.heis_sweeps
            assert np.array_equal(got[k], run(OM, 100 + k, 5, 40, states[k])), k


def test_graph_energy_follows_reference_total_energy(vg, orc, ctx):
    """E and |M| per sweep against the oracle's sequential restatement's
    total_energy of the downloaded state over the caller's matrix (row order,
    duplicates and diagonal as given): Compound(Exchange, Zeeman) and
    Exchange alone."""
    rng = np.random.default_rng(3)
    m = random_csr(rng, 80, 0.06, symmetric=False, dup=True, diag=True)
    for spin in ("ising", "heisenberg"):
        for terms in (("exchange", "zeeman"), ("exchange",)):
            H = vg.hamiltonian(exchange=1.0, zeeman="zeeman" in terms)
            sk = vg.ISING if spin == "ising" else vg.HEISENBERG
            d = vg.chain(vg.graph_lattice(), sk, vg.RANDOM, H, 5)
            with vg.Batch(ctx, [d], exchange=[m]) as b:
                b.randomize()
                b.set_thermostat(0, 1.2, 0.4)
                e, mg = b.run(4)
                OM = orc.model(spin=spin, terms=terms, temperature=1.2, field=0.4)
                seq = orc.Sequential(None, OM, 1, csr=m)
                seq.set_thermostat(1.2, 0.4)
                st = b.get_ising(0) if spin == "ising" else b.get_heisenberg(0)
                seq.set_state(st)
                tol = 1e-9 if spin == "ising" else 1e-5
                assert e[0, -1] == pytest.approx(seq.total_energy(), rel=tol, abs=tol), terms
                assert mg[0, -1] == pytest.approx(seq.magnetization(), rel=tol, abs=tol)
                seq.close()


def test_lattice_matrix_reproduces_lattice_equilibrium(vg, orc, ctx):
    """sc 6^3 Ising at T = 5 with h = 0.1 given as its exchange matrix: 64
    graph chains on the GPU against the reference's sequential Metropolis
    over the same matrix (E/N, |M|/N, C, chi within 3 sigma)."""
    csr = lattice_csr(orc, "sc", (6, 6, 6), (1, 1, 1))
    n, T, h = 216, 5.0, 0.1
    H = vg.hamiltonian(exchange=1.0, zeeman=True)
    descs = [vg.chain(vg.graph_lattice(), vg.ISING, vg.FLIP, H, s) for s in range(64)]
    with vg.Batch(ctx, descs, exchange=[csr] * 64) as b:
        assert b.colours(0) == 2
        b.randomize()
        for i in range(64):
            b.set_thermostat(i, T, h)
        b.set_graph(True)
        b.run(1000, observe=False)
        e, m = b.run(5000)
    obs = (e.mean(1) / n, m.mean(1) / n, e.var(1) / (n * T * T), m.var(1) / (n * T))
    gpu = [(float(o.mean()), float(o.std(ddof=1) / 8.0)) for o in obs]
    OM = orc.model(terms=("exchange", "zeeman"), temperature=T, field=h)
    seq = orc.Sequential(None, OM, 77, csr=csr)
    seq.set_thermostat(T, h)
    seq.set_state(np.ones(n, dtype=np.int8))
    seq.sweeps(2000, observe=False)
    es, ms = seq.sweeps(60000)
    seq.close()
    from _stats import binned_fn
    cpu = [binned(es / n), binned(ms / n), binned_fn(lambda x: x.var() / (n * T * T), [es]),
           binned_fn(lambda x: x.var() / (n * T), [ms])]
    for lab, (a, ea), (c, ec) in zip(["E/N", "|M|/N", "C", "chi"], gpu, cpu):
        z = z_score(a, ea, c, ec)
        print(f"graph sc6 {lab}: engine {a:.5f} +- {ea:.5f} sequential {c:.5f} +- {ec:.5f} z={z:.2f}")
        assert z < 3.0, lab


def test_graph_errors(vg, ctx):
    H = vg.hamiltonian(exchange=1.0)
    ok = (np.array([0, 1, 2], dtype=np.uint64), np.array([1, 0], dtype=np.uint64),
          np.array([1.0, 1.0]))
    d = vg.chain(vg.graph_lattice(), vg.ISING, vg.FLIP, H, 1)
    with pytest.raises(vg.VegasError):              # a matrix is needed
        vg.Batch(ctx, [d])
    bad_col = (ok[0], np.array([1, 5], dtype=np.uint64), ok[2])
    with pytest.raises(vg.VegasError) as e:
        vg.Batch(ctx, [d], exchange=[bad_col])
    assert e.value.kind == "InvalidArgument"
    bad_val = (ok[0], ok[1], np.array([1.0, np.nan]))
    with pytest.raises(vg.VegasError):
        vg.Batch(ctx, [d], exchange=[bad_val])
    for spin, prop in ((vg.ISING_BITS, vg.FLIP), (vg.ISING, vg.WOLFF)):
        with pytest.raises(vg.VegasError) as e:
            vg.Batch(ctx, [vg.chain(vg.graph_lattice(), spin, prop, H, 1)], exchange=[ok])
        assert e.value.kind == "NotImplemented"
    with vg.Server(ctx, memory_bytes=1 << 20, max_jobs=1, quantum_sweeps=8) as srv:
        with pytest.raises(vg.VegasError) as e:
            srv.submit(d, vg.relax(10, 2.0))
        assert e.value.kind == "NotImplemented"
