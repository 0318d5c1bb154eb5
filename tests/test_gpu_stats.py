"""Statistical parity: the engine's checkerboard sweeps against the oracle's
restatement of the reference's sequential random-site Metropolis
(integrator.rs:66-138 with the energy.rs terms), on the same lattice,
temperature and Hamiltonian.

<E>/N, <|M|>/N, the specific heat C = Var(E)/(N T^2) and the susceptibility
chi = Var(|M|)/(N T) (the StatSensor formulas, instrument.rs:113-123) must
agree within 3 sigma, sigma combining both sides' errors: the engine's from
the spread of independent chains (one batch, seeds 0..chains-1), the
sequential chain's from 40 bins (jackknife for C and chi).  The two
algorithms differ in update order, so they agree in distribution only; the
seeds are fixed, so every run is the same."""
import numpy as np
import pytest

from _stats import binned, binned_fn, z_score

pytestmark = pytest.mark.gpu

CASES = {
    # name: (spin, dims, pbc, T, h, proposal, terms, K, chains, sweeps, seq_sweeps)
    "ising_sc8_T6": ("ising", (8, 8, 8), (1, 1, 1), 6.0, 0.0, "flip",
                     ("exchange", "zeeman"), 0.0, 64, 6000, 60000),
    # Ordered phase in a field: started in the favoured direction (Down, since
    # the reference's Zeeman site energy is +|h| s.h), as chains started
    # against it stay magnetized the wrong way for thousands of sweeps.
    "ising_sc8_T4_h0.05_random": ("ising", (8, 8, 8), (1, 1, 1), 4.0, 0.05, "random",
                                  ("exchange", "zeeman"), 0.0, 64, 6000, 60000),
    "ising_sq10_T2.5": ("ising", (10, 10, 1), (1, 1, 0), 2.5, 0.0, "flip",
                        ("exchange", "zeeman"), 0.0, 64, 8000, 100000),
    "ising_sc8_T5_exchange_only": ("ising", (8, 8, 8), (1, 1, 1), 5.0, 0.0, "flip",
                                   ("exchange",), 0.0, 64, 6000, 60000),
    "heis_sc8_T1.4": ("heisenberg", (8, 8, 8), (1, 1, 1), 1.4, 0.05, "random",
                      ("exchange", "zeeman", "anisotropy"), 0.1, 64, 4000, 30000),
    "heis_sc8_T2.5": ("heisenberg", (8, 8, 8), (1, 1, 1), 2.5, 0.05, "random",
                      ("exchange", "zeeman", "anisotropy"), 0.1, 64, 4000, 30000),
    # the lowest temperature of the 128^3 configuration: ordered, easy-plane
    # (K > 0 penalises the z axis in the reference's sign), low acceptance
    "heis_sc8_T0.5": ("heisenberg", (8, 8, 8), (1, 1, 1), 0.5, 0.05, "random",
                      ("exchange", "zeeman", "anisotropy"), 0.1, 64, 6000, 60000),
    "ising_bcc5_T8": ("ising", (5, 5, 5), (1, 1, 1), 8.0, 0.1, "flip",
                      ("exchange", "zeeman"), 0.0, 64, 6000, 60000),
    "ising_bcc6x4x3_open_T5": ("ising", (6, 4, 3), (1, 0, 0), 5.0, 0.0, "flip",
                               ("exchange", "zeeman"), 0.0, 64, 6000, 60000),
    "heis_bcc5_T3": ("heisenberg", (5, 5, 5), (1, 1, 1), 3.0, 0.05, "random",
                     ("exchange", "zeeman", "anisotropy"), 0.1, 64, 4000, 30000),
    # fcc: four colour passes per sweep (T_c of the fcc Ising model is ~9.8 J)
    "ising_fcc4_T12": ("ising", (4, 4, 4), (1, 1, 1), 12.0, 0.1, "flip",
                       ("exchange", "zeeman"), 0.0, 64, 6000, 60000),
    "ising_fcc4x3x3_open_T8_random": ("ising", (4, 3, 3), (0, 1, 1), 8.0, 0.0, "random",
                                      ("exchange", "zeeman"), 0.0, 64, 6000, 180000),
    "heis_fcc4_T4": ("heisenberg", (4, 4, 4), (1, 1, 1), 4.0, 0.05, "random",
                     ("exchange", "zeeman", "anisotropy"), 0.1, 64, 4000, 30000),
}
BCC = {"ising_bcc5_T8", "ising_bcc6x4x3_open_T5", "heis_bcc5_T3"}
FCC = {"ising_fcc4_T12", "ising_fcc4x3x3_open_T8_random", "heis_fcc4_T4"}
PER_CELL = {"sc": 1, "bcc": 2, "fcc": 4}


def _observables(e, m, n, T):
    """Per-series <E>/N, <|M|>/N, C, chi (last axis = sweeps)."""
    return (e.mean(-1) / n, m.mean(-1) / n, e.var(-1) / (n * T * T), m.var(-1) / (n * T))


ORDERED_START = {"ising_sc8_T4_h0.05_random"}


def _engine(vg, ctx, kind, spin, dims, pbc, T, h, proposal, terms, K, chains, sweeps, down):
    kinds = {"sc": vg.SC, "bcc": vg.BCC, "fcc": vg.FCC}
    L = vg.lattice(kinds[kind], *dims, pbc=tuple(bool(p) for p in pbc))
    kw = {"zeeman": "zeeman" in terms}
    if "exchange" in terms:
        kw["exchange"] = 1.0
    if "anisotropy" in terms:
        kw["anisotropy"] = K
    H = vg.hamiltonian(**kw)
    sk = vg.ISING if spin == "ising" else vg.HEISENBERG
    descs = [vg.chain(L, sk, vg.FLIP if proposal == "flip" else vg.RANDOM, H, s)
             for s in range(chains)]
    n = int(np.prod(dims)) * PER_CELL[kind]
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        if down:
            for i in range(chains):
                b.set_ising(i, np.full(n, -1, dtype=np.int8))
        for i in range(chains):
            b.set_thermostat(i, T, h)
        b.run(sweeps // 5, observe=False)  # burn-in
        e, m = b.run(sweeps)
    obs = _observables(e, m, n, T)
    return [(float(o.mean()), float(o.std(ddof=1) / np.sqrt(chains))) for o in obs]


def _sequential(orc, kind, spin, dims, pbc, T, h, proposal, terms, K, sweeps, down,
                seed=12345):
    OL = orc.lattice(kind, *dims, pbc=pbc)
    OM = orc.model(spin=spin, proposal=proposal, terms=terms, anisotropy=K, temperature=T,
                   field=h)
    s = orc.Sequential(OL, OM, seed)
    s.set_thermostat(T, h)
    if down:
        s.set_state(np.full(orc.num_sites(OL), -1, dtype=np.int8))
    elif spin == "ising":
        s.set_state(orc.ising_init(OL, seed))
    else:
        s.set_state(orc.heis_init(OL, seed).astype(np.float64))
    s.sweeps(sweeps // 5, observe=False)
    e, m = s.sweeps(sweeps)
    s.close()
    n = orc.num_sites(OL)
    return [binned(e / n), binned(m / n),
            binned_fn(lambda x: x.var() / (n * T * T), [e]),
            binned_fn(lambda x: x.var() / (n * T), [m])]


@pytest.mark.parametrize("name", list(CASES))
def test_checkerboard_agrees_with_sequential_metropolis(vg, orc, ctx, name):
    spin, dims, pbc, T, h, proposal, terms, K, chains, sweeps, seq_sweeps = CASES[name]
    down = name in ORDERED_START
    kind = "bcc" if name in BCC else ("fcc" if name in FCC else "sc")
    gpu = _engine(vg, ctx, kind, spin, dims, pbc, T, h, proposal, terms, K, chains, sweeps,
                  down)
    seq = _sequential(orc, kind, spin, dims, pbc, T, h, proposal, terms, K, seq_sweeps, down)
    labels = ["E/N", "|M|/N", "C", "chi"]
    rows = []
    for lab, (a, ea), (b, eb) in zip(labels, gpu, seq):
        z = z_score(a, ea, b, eb)
        rows.append((lab, a, ea, b, eb, z))
        print(f"{name} {lab:6s} engine {a:.6f} +- {ea:.6f}  sequential {b:.6f} +- {eb:.6f}"
              f"  z={z:.2f}")
    for lab, a, ea, b, eb, z in rows:
        assert z < 3.0, (name, lab, a, ea, b, eb)
        # the errors are small enough for the comparison to mean something
        assert ea < 0.05 * abs(a) + 1e-3 and eb < 0.05 * abs(b) + 1e-3, (name, lab)
