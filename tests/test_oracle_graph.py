"""The oracle's restatement of chains whose exchange is an arbitrary sparse
matrix (Exchange::new(CsMat<f64>), src/energy.rs:170-172, evaluated at
194-213), checked without the engine:

- the f64 exponential against libm;
- the greedy colouring: proper on the symmetrised pattern, and each site's
  colour is the smallest one its lower-numbered neighbours leave free;
- the per-sweep bond sum equals sum_{i != j} J_ij s_i.s_j of the state, for
  symmetric and non-symmetric matrices;
- equilibrium averages of the colour-class sweeps against exact enumeration
  of small random graphs (mixed-sign couplings, a field, duplicate and
  diagonal entries), and the reference's sequential Metropolis over the same
  matrix against the same enumeration;
- Heisenberg: the two-spin closed form <s1.s2> = coth(J/T) - T/J and the
  independent-spin Langevin form in a field (an empty matrix)."""
import itertools
import math

import numpy as np
import pytest

from _stats import binned, z_score


def random_graph(rng, n, p, jscale=1.0, symmetric=True, dup=False, diag=False):
    """CSR (row_ptr, col, val) of a random coupling matrix."""
    J = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < p:
                J[i, j] = rng.normal() * jscale
                J[j, i] = J[i, j] if symmetric else rng.normal() * jscale
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(n):
            if J[i, j] != 0.0:
                if dup:  # split the coupling into two entries of the row
                    rows += [i, i]
                    cols += [j, j]
                    vals += [0.25 * J[i, j], 0.75 * J[i, j]]
                else:
                    rows.append(i)
                    cols.append(j)
                    vals.append(J[i, j])
        if diag:
            rows.append(i)
            cols.append(i)
            vals.append(0.5)
    order = np.argsort(np.array(rows), kind="stable")
    rows = np.array(rows)[order]
    cols = np.array(cols, dtype=np.uint64)[order]
    vals = np.array(vals)[order]
    row_ptr = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(row_ptr, rows + 1, 1)
    return np.cumsum(row_ptr).astype(np.uint64), cols, vals


def dense(row_ptr, col, val, n):
    J = np.zeros((n, n))
    for i in range(n):
        for k in range(int(row_ptr[i]), int(row_ptr[i + 1])):
            J[i, int(col[k])] += val[k]
    return J


def test_exp_f64_against_libm(orc):
    xs = np.concatenate([-np.linspace(0.0, 708.0, 20001), -np.geomspace(1e-12, 1.0, 500)])
    rel = max(abs(orc.exp_f64_nonpos(x) - math.exp(x)) / math.exp(x) for x in xs)
    assert rel < 1e-15
    assert orc.exp_f64_nonpos(0.0) == 1.0
    assert orc.exp_f64_nonpos(-709.0) == 0.0


@pytest.mark.parametrize("seed", range(6))
def test_greedy_colouring(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 60))
    csr = random_graph(rng, n, float(rng.uniform(0.05, 0.5)), symmetric=seed % 2 == 0,
                       dup=seed % 3 == 0, diag=seed % 2 == 1)
    G = orc.Graph(*csr)
    col = G.colouring()
    J = dense(*csr, n)
    adj = (J != 0) | (J.T != 0)
    np.fill_diagonal(adj, False)
    for i in range(n):
        used = {int(col[j]) for j in range(i) if adj[i, j]}
        want = 0
        while want in used:
            want += 1
        assert col[i] == want, i
        for j in range(n):
            if adj[i, j]:
                assert col[i] != col[j]
    assert G.colours() == int(col.max()) + 1
    assert G.diag() == pytest.approx(float(np.trace(J)))


def test_lattice_matrix_is_two_coloured(orc):
    """The sc lattice's exchange matrix (as the sequential restatement
    builds it) is bipartite and greedily two-coloured by site parity."""
    L = orc.lattice("sc", 4, 6, 2)
    n = orc.num_sites(L)
    rows = [sorted(j for j in orc.neighbours(L, i) if j >= 0) for i in range(n)]
    row_ptr = np.cumsum([0] + [len(r) for r in rows]).astype(np.uint64)
    col = np.array([j for r in rows for j in r], dtype=np.uint64)
    G = orc.Graph(row_ptr, col, np.ones(len(col)))
    assert G.colours() == 2
    x = np.arange(n) % 4
    y = np.arange(n) // 4 % 6
    z = np.arange(n) // 24
    assert np.array_equal(G.colouring(), (x + y + z) % 2)


@pytest.mark.parametrize("symmetric", [True, False])
def test_bond_sum_equals_state_sum(orc, symmetric):
    rng = np.random.default_rng(3)
    n = 40
    csr = random_graph(rng, n, 0.15, symmetric=symmetric, dup=True, diag=True)
    J = dense(*csr, n)
    Joff = J - np.diag(np.diag(J))
    G = orc.Graph(*csr)
    M = orc.model(terms=("exchange", "zeeman"), temperature=1.3, field=0.2)
    s = G.ising_init(5)
    for t in range(6):
        sums = G.ising_sweeps(M, 5, t, 1, s)[0]
        sv = s.astype(np.float64)
        assert sums[3] == int(s.sum())
        assert sums[0] * 2.0 ** -27 == pytest.approx(sv @ Joff @ sv, abs=n * 2.0 ** -27)
    Mh = orc.model(spin="heisenberg", proposal="random", terms=("exchange", "anisotropy"),
                   anisotropy=0.3, temperature=0.9)
    x = G.heis_init(6)
    for t in range(4):
        sums = G.heis_sweeps(Mh, 6, t, 1, x)[0]
        xv = x.astype(np.float64)
        bond = float(np.einsum("ia,ij,ja->", xv, Joff, xv))
        assert sums[0] * 2.0 ** -27 == pytest.approx(bond, abs=1e-4 * (1 + abs(bond)))
        assert sums[3] == sum(int(round(float(np.float32(v) * np.float32(2 ** 27))))
                              for v in x[:, 2])


def ising_exact_matrix(J, h, T):
    """Exact <E>, <|M|> of the Compound(Exchange(J), Zeeman) chain with a
    symmetric J: the dynamics sample exp(-H/T), H = -sum_{i<j} J_ij s_i s_j
    - J_ii/2... (diagonal constant) + |h| sum s; E_obs = sum_i e_i."""
    n = len(J)
    Joff = J - np.diag(np.diag(J))
    states = np.array(list(itertools.product([-1.0, 1.0], repeat=n)))
    pair = np.einsum("si,ij,sj->s", states, Joff, states)  # sum over ordered pairs
    m = states.sum(axis=1)
    H = -0.5 * pair + h * m
    e_obs = -pair - np.trace(J) + h * m
    w = np.exp(-(H - H.min()) / T)
    Z = w.sum()
    return (w * e_obs).sum() / Z, (w * np.abs(m)).sum() / Z


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_ising_graph_matches_exact_enumeration(orc, seed):
    rng = np.random.default_rng(seed)
    n = 9
    csr = random_graph(rng, n, 0.45, jscale=0.8, dup=seed == 12, diag=seed == 13)
    J = dense(*csr, n)
    T, h = 1.1, 0.3
    ee, em = ising_exact_matrix(J, h, T)
    G = orc.Graph(*csr)
    for proposal in ("flip", "random"):
        M = orc.model(terms=("exchange", "zeeman"), temperature=T, field=h, proposal=proposal)
        s = G.ising_init(seed)
        sums = G.ising_sweeps(M, seed, 0, 120000, s)[500:]
        e = -sums[:, 0] * 2.0 ** -27 - G.diag() + h * sums[:, 3]
        me, se = binned(e)
        mm, sm = binned(np.abs(sums[:, 3]).astype(np.float64))
        assert z_score(me, se, ee, 0.0) < 4, (proposal, me, se, ee)
        assert z_score(mm, sm, em, 0.0) < 4, (proposal, mm, sm, em)
    # the reference's sequential Metropolis over the same matrix
    M = orc.model(terms=("exchange", "zeeman"), temperature=T, field=h)
    seq = orc.Sequential(None, M, seed, csr=csr)
    seq.set_thermostat(T, h)
    seq.set_state(np.ones(n, dtype=np.int8))
    seq.sweeps(500, observe=False)
    e, m = seq.sweeps(120000)
    seq.close()
    me, se = binned(e)
    mm, sm = binned(m)
    assert z_score(me, se, ee, 0.0) < 4 and z_score(mm, sm, em, 0.0) < 4


def _langevin(a):
    return 1.0 / math.tanh(a) - 1.0 / a


def test_heisenberg_graph_closed_forms(orc):
    # dimer with an antiferromagnetic coupling given as two unequal entries
    row_ptr = np.array([0, 2, 3], dtype=np.uint64)
    col = np.array([1, 1, 0], dtype=np.uint64)
    val = np.array([-0.3, -0.4, -0.7])
    G = orc.Graph(row_ptr, col, val)
    assert G.colours() == 2
    T = 0.6
    for proposal in ("random", "flip"):
        M = orc.model(spin="heisenberg", proposal=proposal, terms=("exchange",), temperature=T)
        x = G.heis_init(4)
        sums = G.heis_sweeps(M, 4, 0, 150000, x)[200:]
        c = sums[:, 0] * 2.0 ** -27 / (2 * -0.7)    # bond = 2 J s1.s2
        mean, err = binned(c)
        if proposal == "random":
            assert z_score(mean, err, _langevin(-0.7 / T), 0.0) < 4, (mean, err)
    # independent spins in a field: <s.h> = -L(|h|/T) (site energy +|h| s.h)
    n = 64
    G0 = orc.Graph(np.zeros(n + 1, dtype=np.uint64), np.zeros(0, dtype=np.uint64),
                   np.zeros(0))
    assert G0.colours() == 1
    h, T = 0.7, 0.9
    M = orc.model(spin="heisenberg", proposal="random", terms=("exchange", "zeeman"),
                  temperature=T, field=h)
    x = G0.heis_init(2)
    sums = G0.heis_sweeps(M, 2, 0, 4000, x)[100:]
    assert not sums[:, 0].any()
    mz = sums[:, 3] * 2.0 ** -27 / n
    mean, err = binned(mz)
    assert z_score(mean, err, -_langevin(h / T), 0.0) < 4
