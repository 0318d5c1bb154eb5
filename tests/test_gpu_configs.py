"""The GPU configurations at their real sizes, checked against the oracle.

- Heisenberg sc 128^3 x 3 (J = 1, K = 0.1 along z, B = 0.05 along z,
  uniform-sphere proposals) at T = 0.5, 1.4, 2.5: bit-exact sums and spins
  against the oracle's checkerboard after two sweeps; unit norms and the
  energy against the reference's site-by-site total_energy of the downloaded
  state (1e-6 relative).
- The 512-chain mixed-size Ising batch (sc L^3, L in {16..128}, T in
  [3.5, 5.5], h in [-0.1, 0.1], seeds 0..511): every chain's reported sums
  equal the ones recomputed from its downloaded spins, and two chains per
  side length (the 128^3 ones included) equal the oracle bit for bit.
- The served job mix (60% Ising Relax, 30% Heisenberg CoolDown, 10%
  Heisenberg HysteresisLoop, the stage schedules of the served stream with
  fewer sweeps per stage): every job's results equal alone and in the mix,
  and sampled jobs of each kind equal the oracle's run of the same program.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from _program import oracle_lattice, oracle_model, oracle_program, observables

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import bench  # noqa: E402  (the workload definitions: batched set and served mix)

pytestmark = pytest.mark.gpu


def _bond_sum_sc(s):
    """sum over the bonds of s_i s_j (each bond once) of a periodic sc lattice,
    s indexed [z][y][x]; works for scalar (int) and vector (..., 3) spins."""
    tot = 0.0 if s.dtype.kind == "f" else 0
    for a in range(3):
        prod = s * np.roll(s, 1, axis=a)
        tot += prod.sum()
    return tot


def test_heisenberg_128_cube_three_temperatures(vg, orc, ctx):
    side, temps = 128, (0.5, 1.4, 2.5)
    L = vg.lattice(vg.SC, side, side, side)
    H = vg.hamiltonian(exchange=1.0, zeeman=True, anisotropy=0.1)
    descs = [vg.chain(L, vg.HEISENBERG, vg.RANDOM, H, 11 + k) for k in range(3)]
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        for k, T in enumerate(temps):
            b.set_thermostat(k, T, 0.05, (0.0, 0.0, 1.0))
        sums = b.run_sums(2)
        e, m = b.run(1)
        for k, T in enumerate(temps):
            d = descs[k]
            OL = oracle_lattice(orc, d.lattice)
            OM = oracle_model(orc, d, T, 0.05, (0.0, 0.0, 1.0))
            xyz = orc.heis_init(OL, d.seed)
            want = orc.heis_sweeps(OL, OM, d.seed, 0, 2, xyz)
            print(f"heis 128^3 T={T}: sums after 2 sweeps {sums[k, -1].tolist()}", flush=True)
            assert np.array_equal(sums[k], want), T
            want3 = orc.heis_sweeps(OL, OM, d.seed, 2, 1, xyz)
            got = b.get_heisenberg(k)
            assert np.array_equal(got, xyz.astype(np.float64)), T
            norm = np.linalg.norm(got, axis=1)
            assert np.abs(norm - 1.0).max() < 1e-6
            # E and |M| after sweep 3 from the sums, against the reference's
            # total_energy (Compound: sum_i e_i) of the downloaded state.
            ew, mw = observables(d, side ** 3, want3[0], 0.05, (0.0, 0.0, 1.0))
            assert e[k, 0] == ew and m[k, 0] == mw
            s = got.reshape(side, side, side, 3)
            e_ref = (-2.0 * _bond_sum_sc(s) + 0.05 * s[..., 2].sum()
                     + 0.1 * (s[..., 2] ** 2).sum())
            m_ref = np.linalg.norm(s.reshape(-1, 3).sum(axis=0))
            assert e[k, 0] == pytest.approx(e_ref, rel=1e-6), T
            assert m[k, 0] == pytest.approx(m_ref, rel=1e-6, abs=1e-6), T


def test_batched_512_chain_set(vg, orc, ctx):
    sides, descs, therm = bench.batched_chains(vg, 512)
    nsw = 3
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        for i, (T, h, hd) in enumerate(therm):
            b.set_thermostat(i, T, h, hd)
        tile = b.autotune(2)            # restores spins and counters
        b.set_graph(True)
        sums = b.run_sums(nsw)
        print(f"batched 512: tile {tile}, {sum(int(L) ** 3 for L in sides)} sites", flush=True)
        # every chain: reported sums == sums of its downloaded state
        for i, L in enumerate(sides):
            L = int(L)
            s = b.get_ising(i).astype(np.int64).reshape(L, L, L)
            assert sums[i, -1, 3] == s.sum(), i
            assert sums[i, -1, 0] == _bond_sum_sc(s), i
            assert b.sweep_count(i) == nsw
        # two chains per side length against the oracle, bit for bit
        picked = []
        for Lv in sorted(set(int(x) for x in sides)):
            picked += [i for i, x in enumerate(sides) if int(x) == Lv][:2]
        for i in picked:
            d = descs[i]
            T, h, hd = therm[i]
            OL = oracle_lattice(orc, d.lattice)
            OM = oracle_model(orc, d, T, h, hd)
            spins = orc.ising_init(OL, d.seed)
            want = orc.ising_sweeps(OL, OM, d.seed, 0, nsw, spins)
            print(f"batched chain {i}: L={int(sides[i])} T={T:.4f} h={h:.4f} dir={hd[2]:+.0f}",
                  flush=True)
            assert np.array_equal(sums[i], want), i
            assert np.array_equal(b.get_ising(i), spins), i
    assert any(int(sides[i]) == 128 for i in picked)


def _scaled_served_jobs(vg, n_jobs, relax_sweeps=40, cool_sweeps=6, loop_sweeps=2):
    """The served mix (bench.served_jobs: kinds, lattices, temperatures,
    field schedules, seeds) with fewer sweeps per stage."""
    out = []
    for kind, d, stages, _, _ in bench.served_jobs(vg, n_jobs):
        if kind == "ising-relax":
            st = vg.relax(relax_sweeps, stages[0].temperature)
        elif kind == "heis-cooldown":
            st = vg.cool_down(3.0, 0.1, 0.1, relax=0, steps=cool_sweeps)
            assert len(st) == len(stages) == 29
        else:
            st = vg.hysteresis_loop(steps=loop_sweeps, relax=0, temperature=1.0,
                                    max_field=0.5, field_step=0.05)
            assert len(st) == len(stages)
        out.append((kind, d, st))
    return out


def test_served_mix_equals_oracle_alone_and_in_mix(vg, orc, ctx):
    jobs = _scaled_served_jobs(vg, 40)
    kinds = [k for k, _, _ in jobs]
    assert {"ising-relax", "heis-cooldown", "heis-hysteresis"} <= set(kinds)
    # In the mix: all submitted at once into 16 slots and a 16 MiB pool (the
    # jobs' spins take about 30 MiB together), so jobs queue, are admitted as
    # others leave and share sweep blocks with different quantum cuts.
    pool = 16 << 20
    with vg.Server(ctx, memory_bytes=pool, max_jobs=16, quantum_sweeps=25) as srv:
        ids = [srv.submit(d, st) for _, d, st in jobs]
        mixed, last = [], []
        for jid in ids:
            s = srv.wait(jid, 300.0)
            assert s.state == vg.JOB_DONE, s.state
            mixed.append(srv.results(jid))
            last.append((s.last_energy, s.last_magnetization))
        stats = srv.stats()
        assert stats.peak_jobs > 1 and stats.completed == len(jobs)
        assert stats.pool_peak <= pool
    # alone: one job at a time
    with vg.Server(ctx, memory_bytes=pool, max_jobs=1, quantum_sweeps=64) as srv:
        for k, (_, d, st) in enumerate(jobs):
            jid = srv.submit(d, st)
            s = srv.wait(jid, 300.0)
            assert s.state == vg.JOB_DONE
            assert np.array_equal(srv.results(jid), mixed[k]), k
            assert (s.last_energy, s.last_magnetization) == last[k], k
            srv.release(jid)
    # Sampled jobs of each kind against the oracle's run of the same program:
    # the stage lines and the last sweep's E and |M| (a Relax job measures
    # nothing, so its trajectory is pinned through them).
    for kind in ("ising-relax", "heis-cooldown", "heis-hysteresis"):
        for k in [i for i, x in enumerate(kinds) if x == kind][:2]:
            _, d, st = jobs[k]
            print(f"served job {k}: {kind} L={d.lattice.lx} seed={d.seed}", flush=True)
            want, want_last = oracle_program(orc, d, st)
            assert np.array_equal(mixed[k], want), k
            assert last[k] == want_last, k
