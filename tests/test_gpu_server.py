"""Job server (continuous batching): results are bit-identical alone and in
the mix, equal to the same program driven through a plain batch, and
admission, queueing and cancellation behave as documented."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(vg):
    """A mix of kinds, sizes and programs (short stages)."""
    J = []
    ising = vg.hamiltonian(exchange=1.0, zeeman=True)
    heis = vg.hamiltonian(exchange=1.0, zeeman=True, anisotropy=0.1)
    for k, L in enumerate([16, 24, 32, 10, 40, 18]):
        d = vg.chain(vg.lattice(vg.SC, L, L, L), vg.ISING, vg.FLIP, ising, 100 + k)
        J.append((d, vg.relax(steps=150 + 37 * k, temperature=2.0 + 0.5 * k)))
    for k, L in enumerate([8, 12, 16]):
        d = vg.chain(vg.lattice(vg.SC, L, L, L), vg.HEISENBERG, vg.RANDOM, heis, 200 + k)
        J.append((d, vg.cool_down(3.0, 2.0, 0.25, relax=20, steps=60 + 10 * k)))
    d = vg.chain(vg.lattice(vg.BCC, 6, 6, 6), vg.ISING, vg.FLIP, ising, 300)
    J.append((d, vg.hysteresis_loop(steps=40, relax=10, temperature=4.0, max_field=0.3,
                                     field_step=0.1)))
    d = vg.chain(vg.lattice(vg.BCC, 5, 5, 5), vg.HEISENBERG, vg.RANDOM, heis, 301)
    J.append((d, vg.cool_down(2.5, 1.5, 0.5, relax=0, steps=70)))
    d = vg.chain(vg.lattice(vg.SC, 12, 12, 12), vg.ISING, vg.RANDOM, ising, 302)
    J.append((d, vg.cool_down(5.0, 4.0, 0.5, relax=33, steps=101, field=0.1)))
    d = vg.chain(vg.lattice(vg.BCC, 20, 6, 6), vg.ISING_BITS, vg.FLIP, ising, 303)
    J.append((d, vg.relax(steps=90, temperature=5.0)))
    return J


def test_alone_equals_in_mix(vg, ctx):
    jobs = _jobs(vg)
    alone = []
    with vg.Server(ctx, memory_bytes=64 << 20, max_jobs=8, quantum_sweeps=64) as srv:
        for d, stages in jobs:
            jid = srv.submit(d, stages)
            st = srv.wait(jid, 120.0)
            assert st.state == vg.JOB_DONE
            alone.append(srv.results(jid))
    # All at once, with few slots and a small pool: jobs queue, get admitted
    # as others finish, and share sweep blocks with different quantum cuts;
    # then with a per-sweep budget that makes the shortest jobs run first and
    # the others sit quanta out.
    pool = 400 << 10  # the jobs need ~700 KiB together
    for budget in (0, 40000):
        with vg.Server(ctx, memory_bytes=pool, max_jobs=5, quantum_sweeps=48,
                       sweep_budget=budget) as srv:
            ids = [srv.submit(d, stages) for d, stages in jobs]
            mixed = []
            for jid in ids:
                st = srv.wait(jid, 120.0)
                assert st.state == vg.JOB_DONE, st.state
                mixed.append(srv.results(jid))
            stats = srv.stats()
            assert stats.peak_jobs > 1            # work overlapped on the device
            assert stats.pool_peak <= pool
            assert stats.completed == len(jobs)
        for k, (a, b) in enumerate(zip(alone, mixed)):
            assert np.array_equal(a, b), (budget, k)


def test_graph_replay_equals_plain_quanta(vg, ctx):
    """With use_graph the server replays each quantum from a captured
    hipGraph while the job set is unchanged (recaptured when it changes or
    the quantum length differs): every job's results equal plain launches."""
    jobs = _jobs(vg)
    runs = []
    for graph in (False, True):
        with vg.Server(ctx, memory_bytes=400 << 10, max_jobs=5, quantum_sweeps=48,
                       use_graph=graph) as srv:
            ids = [srv.submit(d, stages) for d, stages in jobs]
            out = []
            for jid in ids:
                st = srv.wait(jid, 120.0)
                assert st.state == vg.JOB_DONE, st.state
                out.append(srv.results(jid))
            runs.append(out)
    for k, (a, b) in enumerate(zip(*runs)):
        assert np.array_equal(a, b), k


def test_server_equals_batch_program(vg, ctx):
    """The server's stage results equal the same program driven through a
    one-chain batch with the StatSensor formulas, accumulated in sweep order."""
    heis = vg.hamiltonian(exchange=1.0, zeeman=True, anisotropy=0.1)
    d = vg.chain(vg.lattice(vg.SC, 10, 10, 10), vg.HEISENBERG, vg.RANDOM, heis, 77)
    stages = vg.cool_down(2.0, 1.0, 0.5, relax=25, steps=90, field=0.05)
    with vg.Server(ctx, memory_bytes=16 << 20, max_jobs=4, quantum_sweeps=32) as srv:
        jid = srv.submit(d, stages)
        assert srv.wait(jid, 60.0).state == vg.JOB_DONE
        got = srv.results(jid)
    want = []
    n = 1000
    with vg.Batch(ctx, [d]) as b:
        b.randomize()
        for s in stages:
            b.set_thermostat(0, s.temperature, s.field, tuple(s.field_dir))
            b.run(s.relax, observe=False)
            e, m = b.run(s.measure)
            se = se2 = sm = sm2 = sm4 = 0.0
            for x, y in zip(e[0], m[0]):
                se += x
                se2 += x * x
                sm += y
                sm2 += y * y
                sm4 += y * y * y * y
            c = float(s.measure)
            T = s.temperature
            me, mm = se / c, sm / c
            want.append([T, abs(s.field), me, (se2 / c - me * me) / (n * (T * T)), mm,
                         (sm2 / c - mm * mm) / (n * T), 1.0 - (sm4 / c) / (3.0 * ((sm2 / c) ** 2)),
                         c])
    assert np.array_equal(got, np.array(want))


def test_admission_queue_cancel(vg, ctx):
    ising = vg.hamiltonian(exchange=1.0, zeeman=True)
    big = vg.chain(vg.lattice(vg.SC, 128, 128, 128), vg.ISING, vg.FLIP, ising, 1)
    small = vg.chain(vg.lattice(vg.SC, 16, 16, 16), vg.ISING, vg.FLIP, ising, 2)
    with vg.Server(ctx, memory_bytes=1 << 20, max_jobs=1, quantum_sweeps=16) as srv:
        with pytest.raises(vg.VegasError) as e:
            srv.submit(big, vg.relax(10, 3.0))      # 2 MiB of spins, 1 MiB pool
        assert e.value.kind == "Admission"
        with pytest.raises(vg.VegasError) as e:
            srv.submit(small, [])
        assert e.value.kind == "NoSteps"
        a = srv.submit(small, vg.relax(3000, 3.0))
        b = srv.submit(small, vg.relax(10, 3.0))    # waits: one slot
        c = srv.submit(small, vg.relax(10, 3.0))
        srv.cancel(c)
        assert srv.wait(b, 60.0).state == vg.JOB_DONE
        assert srv.wait(c, 60.0).state == vg.JOB_CANCELLED
        sa, sb = srv.status(a), srv.status(b)
        assert sa.state == vg.JOB_DONE and sb.start_s >= sa.end_s  # FIFO, one slot
        with pytest.raises(vg.VegasError) as e:
            srv.results(c)
        assert e.value.kind == "Cancelled"
        with pytest.raises(vg.VegasError) as e:
            srv.status(12345)
        assert e.value.kind == "NotFound"
        srv.release(b)
        with pytest.raises(vg.VegasError):
            srv.status(b)


def test_concurrent_submitters(vg, ctx):
    """Jobs submitted from several threads while others run all finish, each
    with the result it has alone."""
    ising = vg.hamiltonian(exchange=1.0, zeeman=True)
    descs = [vg.chain(vg.lattice(vg.SC, 16, 16, 16), vg.ISING, vg.FLIP, ising, 500 + k)
             for k in range(12)]
    progs = [vg.relax(200 + 13 * k, 2.5 + 0.1 * k) for k in range(12)]
    with vg.Server(ctx, memory_bytes=32 << 20, max_jobs=64, quantum_sweeps=40) as srv:
        ref = []
        for d, p in zip(descs, progs):
            jid = srv.submit(d, p + vg.cool_down(2.0, 2.0, 0.1, relax=0, steps=50))
            srv.wait(jid, 60.0)
            ref.append(srv.results(jid))
        ids = [None] * 12

        def worker(k0):
            for k in range(k0, 12, 3):
                ids[k] = srv.submit(descs[k], progs[k] + vg.cool_down(2.0, 2.0, 0.1, relax=0,
                                                                       steps=50))
        th = [threading.Thread(target=worker, args=(t,)) for t in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for k in range(12):
            assert srv.wait(ids[k], 60.0).state == vg.JOB_DONE
            assert np.array_equal(srv.results(ids[k]), ref[k]), k
