"""The reference's CPU-runnable configuration on the GPU: the CLI program on
the 10x10x1 sample (x, y periodic, z open; Compound(Exchange J=1, Zeeman),
h = 0; Ising spins from seed 42; flip proposals): Relax 1000 sweeps at
T = 4.0, then CoolDown 4.0 -> 0.1 by 0.1 (39 temperatures) with 1000 relax +
20000 measured sweeps each -- 819,000 sweeps.

Above T = 2.5 every stage's <E>/N, <|M|>/N, C and chi agree with the
sequential Metropolis restatement run the same way within 4 sigma (60
comparisons; binned errors).  Below, single 2D chains may freeze into
stripe states, so there the energies are only checked against the ground
state (E/N = -4, Compound double counting) and the one-stripe-pair bound."""
import numpy as np
import pytest

from _stats import binned, binned_fn, z_score

pytestmark = pytest.mark.gpu


def _schedule(vg):
    return vg.relax(1000, 4.0) + vg.cool_down(4.0, 0.1, 0.1, relax=1000, steps=20000)


def test_config1_cli_program(vg, orc, ctx):
    stages = _schedule(vg)
    assert len(stages) == 40
    L = vg.lattice(vg.SC, 10, 10, 1, pbc=(True, True, False))
    d = vg.chain(L, vg.ISING, vg.FLIP, vg.hamiltonian(exchange=1.0, zeeman=True), 42)
    n = 100
    gpu = []
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        b.set_graph(True)
        for st in stages:
            b.set_thermostat(0, st.temperature, st.field)
            if st.relax:
                b.run(st.relax, observe=False)
            if st.measure:
                e, m = b.run(st.measure)
                gpu.append((st.temperature, e[0], m[0]))
    OL = orc.lattice("sc", 10, 10, 1, pbc=(1, 1, 0))
    OM = orc.model(terms=("exchange", "zeeman"), temperature=4.0)
    seq = orc.Sequential(OL, OM, 42)
    seq.set_state(orc.ising_init(OL, 42))
    cpu = []
    for st in stages:
        seq.set_thermostat(st.temperature, st.field)
        if st.relax:
            seq.sweeps(st.relax, observe=False)
        if st.measure:
            e, m = seq.sweeps(st.measure)
            cpu.append((st.temperature, e, m))
    seq.close()
    assert len(gpu) == len(cpu) == 39
    compared = 0
    for (T, eg, mg), (_, ec, mc) in zip(gpu, cpu):
        if T >= 2.5:
            for f in (lambda e, m: e / n, lambda e, m: m / n):
                a, ea = binned(f(eg, mg))
                c, ec_ = binned(f(ec, mc))
                assert z_score(a, ea, c, ec_) < 4.0, (T, a, ea, c, ec_)
                compared += 1
            Cg = binned_fn(lambda x: x.var() / (n * T * T), [eg])
            Cc = binned_fn(lambda x: x.var() / (n * T * T), [ec])
            assert z_score(*Cg, *Cc) < 4.0, (T, Cg, Cc)
            Xg = binned_fn(lambda x: x.var() / (n * T), [mg])
            Xc = binned_fn(lambda x: x.var() / (n * T), [mc])
            assert z_score(*Xg, *Xc) < 4.0, (T, Xg, Xc)
            compared += 2
        else:
            for e in (eg, ec):
                assert e.min() / n >= -4.0 - 1e-12
            if T <= 1.0:
                assert eg.mean() / n <= -3.0 and ec.mean() / n <= -3.0, T
    # 4.0 down to 2.6 (the f64 steps pass 2.5 at 2.4999999999999996): 15 stages
    assert compared == 4 * sum(1 for T, *_ in gpu if T >= 2.5) == 60
