"""fp32 Heisenberg engine vs the CPU checkerboard oracle.  The tolerance the
engine must meet is 1e-5 per spin component and 1e-6 relative on energies;
both paths perform the same IEEE fp32 operations, so the check here is
exact equality, which implies both."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COMPONENT_TOL = 1e-5
ENERGY_RTOL = 1e-6


def make_pair(vg, orc, dims, pbc=(1, 1, 1), T=1.4, h=0.05, hdir=(0.0, 0.0, 1.0),
              proposal="random", K=0.1, axis=(0.0, 0.0, 1.0), seed=3,
              terms=("exchange", "zeeman", "anisotropy")):
    L = vg.lattice(vg.SC, *dims, pbc=tuple(bool(p) for p in pbc))
    H = vg.hamiltonian(exchange=1.0 if "exchange" in terms else None,
                       zeeman="zeeman" in terms,
                       anisotropy=K if "anisotropy" in terms else None, axis=axis,
                       gauge=0.5 if "gauge" in terms else None)
    d = vg.chain(L, vg.HEISENBERG, vg.RANDOM if proposal == "random" else vg.FLIP, H, seed)
    OL = orc.lattice("sc", *dims, pbc=pbc)
    OM = orc.model(spin="heisenberg", proposal=proposal, terms=terms, anisotropy=K, axis=axis,
                   gauge=0.5, temperature=T, field=h, field_dir=hdir)
    return d, OL, OM


def check_state(got, want):
    got = np.asarray(got)
    want = np.asarray(want, dtype=np.float64)
    assert np.abs(got - want).max() <= COMPONENT_TOL
    assert np.array_equal(got, want)


def test_randomize_matches_oracle(vg, orc, ctx):
    for dims, pbc in [((16, 16, 16), (1, 1, 1)), ((10, 10, 1), (1, 1, 0))]:
        d, OL, _ = make_pair(vg, orc, dims, pbc, seed=11)
        with vg.Batch(ctx, [d]) as b:
            b.randomize()
            check_state(b.get_heisenberg(0), orc.heis_init(OL, 11))


def test_config_shape_bit_exact_after_every_sweep(vg, orc, ctx):
    """The 128^3 configuration's Hamiltonian (J = 1, K = 0.1 along z, B = 0.05
    along z, uniform-sphere proposals) on a 24^3 lattice at T = 1.4."""
    d, OL, OM = make_pair(vg, orc, (24, 24, 24))
    xyz = orc.heis_init(OL, 3)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        b.set_thermostat(0, 1.4, 0.05, (0.0, 0.0, 1.0))
        for k in range(4):
            got = b.run_sums(1)[0]
            want = orc.heis_sweeps(OL, OM, 3, k, 1, xyz)
            assert np.array_equal(got, want), k
            check_state(b.get_heisenberg(0), xyz)
        got = b.run_sums(6)[0]
        want = orc.heis_sweeps(OL, OM, 3, 4, 6, xyz)
        assert np.array_equal(got, want)
        check_state(b.get_heisenberg(0), xyz)


CASES = [
    # dims, pbc, T, h, hdir, proposal, K, axis, terms
    ((10, 10, 1), (1, 1, 0), 0.7, 0.0, (0, 0, 1), "random", 0.0, (0, 0, 1), ("exchange",)),
    ((2, 4, 6), (1, 1, 1), 1.0, 0.3, (0, 1, 0), "random", 0.5, (1, 0, 0),
     ("exchange", "zeeman", "anisotropy")),
    ((34, 4, 4), (1, 1, 1), 2.5, 0.2, (0.6, 0.0, 0.8), "random", -0.3, (0, 0, 1),
     ("exchange", "zeeman", "anisotropy", "gauge")),
    ((12, 3, 5), (0, 0, 0), 0.5, 0.1, (0, 0, 1), "random", 0.1, (0, 0, 1),
     ("exchange", "zeeman", "anisotropy")),
    ((16, 8, 8), (1, 1, 1), 1.0, 0.4, (0, 0, 1), "flip", 0.0, (0, 0, 1), ("exchange", "zeeman")),
    ((14, 6, 6), (1, 1, 1), 0.05, 0.0, (0, 0, 1), "random", 0.0, (0, 0, 1), ("exchange",)),
]


@pytest.mark.parametrize("dims,pbc,T,h,hdir,proposal,K,axis,terms", CASES)
def test_edge_cases_bit_exact(vg, orc, ctx, dims, pbc, T, h, hdir, proposal, K, axis, terms):
    d, OL, OM = make_pair(vg, orc, dims, pbc, T, h, hdir, proposal, K, axis, 21, terms)
    xyz = orc.heis_init(OL, 21)
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        b.set_thermostat(0, T, h, hdir)
        got = b.run_sums(5)[0]
        want = orc.heis_sweeps(OL, OM, 21, 0, 5, xyz)
        assert np.array_equal(got, want)
        check_state(b.get_heisenberg(0), xyz)


def test_energy_matches_reference_total_energy(vg, orc, ctx):
    """Per-sweep E from the fixed-point sums against the reference's
    site-by-site Compound sum of the downloaded state, within 1e-6 relative."""
    d, OL, OM = make_pair(vg, orc, (12, 12, 12), T=1.4, h=0.05,
                          terms=("exchange", "zeeman", "anisotropy", "gauge"))
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        b.set_thermostat(0, 1.4, 0.05)
        e, m = b.run(3)
        s = orc.Sequential(OL, OM, 1)
        s.set_thermostat(1.4, 0.05)
        s.set_state(b.get_heisenberg(0))
        assert e[0, -1] == pytest.approx(s.total_energy(), rel=ENERGY_RTOL)
        assert m[0, -1] == pytest.approx(s.magnetization(), rel=ENERGY_RTOL)


def test_set_get_and_norm(vg, orc, ctx):
    d, OL, _ = make_pair(vg, orc, (8, 6, 4))
    rng = np.random.default_rng(0)
    v = rng.normal(size=(8 * 6 * 4, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    with vg.Batch(ctx, [d]) as b:
        b.set_heisenberg(0, v)
        got = b.get_heisenberg(0)
        assert np.array_equal(got, v.astype(np.float32).astype(np.float64))
        b.set_thermostat(0, 1.0)
        b.run(3)
        n = np.linalg.norm(b.get_heisenberg(0), axis=1)
        assert np.abs(n - 1).max() < 1e-6
