"""Benchmark of the Metropolis sweep on one MI355X.

Workload (default, `ising-sc256-i8`): three independent Ising chains on the
simple-cubic 256^3 lattice (16.7M sites each), full PBC, J = 1, h = 0, int8
spins, random +-1 start from seed 7, at T = 3.5, 4.5115 and 6.0, advanced by
checkerboard sweeps (flip proposals).  One step = one sweep of all three
chains (3 x 16,777,216 spin-flip attempts), with the per-sweep energy and
|magnetization| of every chain returned to the host.

Prints one JSON line: value = spin-flip attempts/s over the timed steps
(spins resident in HBM), the roofline of the colour-pass kernel (algorithmic
bytes per launch over its mean launch time from HIP events on the engine's
stream), and the CPU baseline (the oracle's restatement of the reference's
sequential MetropolisFlipIntegrator, one core, a bounded sample).

With the default --workload all, the line also carries the other forms of
the path as sub-objects: the same chains bit-packed, one HBM-resident chain
of sc 1024^3 (larger than the Infinity Cache), the 512-chain mixed
batch against per-chain plain launches, Heisenberg sc 128^3 x 3, Ising and
Heisenberg chains on fcc lattices (four colour passes per sweep) and on bcc
lattices, Wolff
cluster steps (batched and single chain, with the oracle's Wolff on one
core), and the served Poisson job stream with p50/p99 latency (--serve-jobs
jobs per rate).
"""
import argparse
import hashlib
import importlib.util
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
COUNTERS = ROOT / "profiles" / "pass_counters.txt"
CSRC = ROOT / "vegas-rs_amd" / "csrc"
# The sources each pass kernel is compiled from (the file and the headers it
# includes): a counter row applies to the build only while they are unchanged.
KERNEL_SOURCES = {
    "ising_sc_pass": ["ising_sc.hip", "ising_swar.hpp", "vg_device.hpp"],
    "ising_bcc_pass": ["ising_bcc.hip", "ising_swar.hpp", "vg_device.hpp"],
    "ising_fcc_pass": ["ising_fcc.hip", "fcc_common.hpp", "ising_swar.hpp", "vg_device.hpp"],
    "heis_sc_pass": ["heis_sc.hip", "heis_common.hpp", "vg_device.hpp"],
    "heis_bcc_pass": ["heis_bcc.hip", "heis_common.hpp", "vg_device.hpp"],
    "heis_fcc_pass": ["heis_fcc.hip", "fcc_common.hpp", "heis_common.hpp", "vg_device.hpp"],
}
SIMDS, CLOCK_GHZ = 1024, 2.4  # MI355X: 256 CUs x 4 SIMDs, peak engine clock
METRIC = "spin-flip attempts/s (Ising sc 256³); batched vs plain; p50/p99 job latency"


def load_engine():
    spec = importlib.util.spec_from_file_location("vegas_hip", ROOT / "vegas-rs_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vegas_hip"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def source_digest(kernel):
    """SHA-1 (12 hex digits) of the sources a pass kernel is compiled from."""
    h = hashlib.sha1()
    for f in KERNEL_SOURCES[kernel.split("<")[0]]:
        h.update((CSRC / f).read_bytes())
    return h.hexdigest()[:12]


def pass_counters(workload, kernels):
    """Counter-measured bytes per dispatch (FETCH_SIZE + WRITE_SIZE, scaled by
    a device copy of known size) and VALU instructions per dispatch of the
    named pass kernels -- the mean over them -- from the table the rocprofv3
    runs of this workload wrote (profiles/pass_counters.txt).  None when the
    table has no row for one of them (another tile, say) or when a kernel's
    sources differ from those the table was measured with (its "# sources"
    line holds their digests)."""
    if not COUNTERS.exists():
        return None
    rows, digests = {}, {}
    for line in COUNTERS.read_text().splitlines():
        if line.startswith("# sources:"):
            digests = dict(kv.split("=") for kv in line[len("# sources:"):].split())
            continue
        if line.startswith("#") or not line.strip():
            continue
        w, k, _, fetch, write, valu, _ = line.split("\t")
        rows[(w, k)] = (float(fetch) + float(write), float(valu))
    got = [rows.get((workload, k)) for k in kernels]
    if any(g is None for g in got):
        return None
    if any(digests.get(k.split("<")[0]) != source_digest(k) for k in kernels):
        return None
    return {"traffic": sum(g[0] for g in got) / len(got),
            "valu": sum(g[1] for g in got) / len(got),
            "source": f"profiles/pass_counters.txt, {workload}: {', '.join(kernels)}"}


def with_counters(roofline, workload, kernels, mads_per_launch=None):
    """Adds the counter-measured traffic (bytes per launch, null when not
    profiled at this tile) and, given the v_mad_u64_u32 count of the launch,
    its VALU ceiling: the time the measured VALU instructions take to issue
    on every SIMD at the peak clock (4 cycles per wave64 instruction, 16 for
    v_mad_u64_u32), as a fraction of the measured launch time."""
    c = pass_counters(workload, kernels)
    roofline["traffic"] = c["traffic"] if c else None
    roofline["traffic_source"] = c["source"] if c else None
    if c and mads_per_launch is not None:
        busy_us = (4.0 * c["valu"] + 12.0 * mads_per_launch) / (SIMDS * CLOCK_GHZ * 1e3)
        roofline["valu_ceiling"] = {
            "valu_instructions_per_launch": c["valu"],
            "mad_u64_per_launch": mads_per_launch,
            "busy_us": busy_us,
            "frac": busy_us / roofline["mean_launch_us"],
            "peak": f"{SIMDS} SIMDs x {CLOCK_GHZ} GHz, one wave64 VALU instruction per 4 cycles "
                    f"(v_mad_u64_u32: 16)",
        }
    return roofline


def ising_mads(attempts_per_launch):
    """v_mad_u64_u32 (wave instructions) of an Ising pass launch with flip
    proposals: one Philox4x32-10 call (20 of them) per 16 sites, 64 lanes per
    wave instruction (tie draws, a few per 10^3 sites, left out).  An upper
    bound for the sc pass: the multiplies of the first two rounds that do not
    depend on the row are computed once per z march, not once per plane."""
    return 20.0 * attempts_per_launch / 1024.0


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(L_side, T, sweeps):
    """The oracle's restatement of the reference's sequential Metropolis flip
    loop (integrator.rs:109-138, Exchange over a CSR matrix, f64, one core):
    one untimed warm-up sweep, then `sweeps` timed ones."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import oracle as orc

    OL = orc.lattice("sc", L_side, L_side, L_side)
    OM = orc.model(terms=("exchange",), temperature=T)
    t0 = time.perf_counter()
    seq = orc.Sequential(OL, OM, 7)
    seq.set_state(orc.ising_init(OL, 7))
    seq.set_thermostat(T)
    build_s = time.perf_counter() - t0
    seq.sweeps(1, observe=False)
    t0 = time.perf_counter()
    seq.sweeps(sweeps, observe=False)
    dt = time.perf_counter() - t0
    n = orc.num_sites(OL)
    seq.close()
    del np
    return {
        "value": n * sweeps / dt,
        "unit": "spin-flip attempts/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{sweeps} timed sweeps ({n * sweeps} attempts) of sc {L_side}^3 Ising, "
                   f"Exchange J=1 over a CSR matrix, flip proposal, T={T}, after a "
                   f"{build_s:.1f} s CSR build and one warm-up sweep; oracle restatement of "
                   f"MetropolisFlipIntegrator::step, 1 thread on {cpu_model()} "
                   f"(nproc {os.cpu_count()})"),
    }


def cpu_threads():
    """Threads for the all-cores CPU lines: the OpenMP setting when given,
    else the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_best_line(L_side, T, sweeps_one, sweeps_all, threads):
    """The build's own CPU checkerboard (the oracle's fast sc path, the same
    draws as the GPU) on one core and on `threads` cores: the best-CPU line."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc

    OL = orc.lattice("sc", L_side, L_side, L_side)
    OM = orc.model(terms=("exchange",), temperature=T)
    spins = orc.ising_init(OL, 7)
    n = orc.num_sites(OL)
    out = {}
    for label, sw, th in (("one_core", sweeps_one, 1), ("all_cores", sweeps_all, threads)):
        t0 = time.perf_counter()
        orc.ising_sweeps_sc_fast(OL, OM, 7, 0, sw, spins, th)
        dt = time.perf_counter() - t0
        out[label] = {"value": n * sw / dt, "cores": th, "sweeps": sw}
    out["unit"] = "spin-flip attempts/s"
    out["sample"] = (f"sc {L_side}^3 Ising checkerboard, Philox draws as on the GPU, T={T}; "
                     f"oracle fast path (OpenMP) on {cpu_model()}")
    return out


def _sequential_worker(job):
    """One chain of the batched set under the oracle's sequential
    restatement for about `seconds`: (attempts, seconds)."""
    side, T, h, seed, seconds = job
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc

    OL = orc.lattice("sc", side, side, side)
    hdir = (0.0, 0.0, 1.0 if h >= 0 else -1.0)
    OM = orc.model(terms=("exchange", "zeeman"), temperature=T, field=abs(h), field_dir=hdir)
    s = orc.Sequential(OL, OM, seed)
    s.set_thermostat(T, abs(h), hdir)
    s.set_state(orc.ising_init(OL, seed))
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        s.sweeps(1, observe=False)
        done += 1
    dt = time.perf_counter() - t0
    s.close()
    return side ** 3 * done, dt


def cpu_batched_aggregate(procs, seconds=3.0, n_chains=512):
    """The batched set on the CPU: min(procs, 512) concurrent processes, each
    running one chain under the oracle's sequential restatement of the
    reference's Metropolis (one core each); the aggregate attempts/s.  Runs
    before the GPU is touched, so the worker processes fork from a process
    without device state."""
    import multiprocessing as mp
    sides, temps, fields = batched_params(n_chains)
    k = min(procs, n_chains)
    jobs = [(int(sides[i]), float(temps[i]), float(fields[i]), i, seconds) for i in range(k)]
    with mp.get_context("fork").Pool(k) as pool:
        res = pool.map(_sequential_worker, jobs)
    span = max(t for _, t in res)
    total = sum(a for a, _ in res)
    return {"value": total / span, "unit": "spin-flip attempts/s", "cores": k, "kind": "port",
            "sample": (f"chains 0..{k - 1} of the batched set, one process each, {seconds:.0f} s of "
                       f"sweeps of the oracle's sequential MetropolisFlipIntegrator restatement "
                       f"on {cpu_model()}")}


def batched_params(n_chains=512, gen_seed=2024):
    """L, T and h of the batched set (numpy PCG64 seeded with gen_seed)."""
    import numpy as np
    rng = np.random.default_rng(gen_seed)
    sides = rng.choice(np.array([16, 24, 32, 48, 64, 96, 128]), size=n_chains)
    temps = rng.uniform(3.5, 5.5, size=n_chains)
    fields = rng.uniform(-0.1, 0.1, size=n_chains)
    return sides, temps, fields


def batched_chains(vg, n_chains=512, gen_seed=2024):
    """The mixed-size batch: n_chains Ising chains on cubic sc lattices, L
    uniform in {16, 24, 32, 48, 64, 96, 128} (full PBC, J = 1), T uniform in
    [3.5, 5.5], h uniform in [-0.1, 0.1] (a negative h is a Down-oriented
    field of magnitude |h|), Compound(Exchange, Zeeman), chain seeds
    0..n_chains-1.  L, T and h come from numpy's PCG64 seeded with gen_seed."""
    sides, temps, fields = batched_params(n_chains, gen_seed)
    H = vg.hamiltonian(exchange=1.0, zeeman=True)
    descs = [vg.chain(vg.lattice(vg.SC, int(L), int(L), int(L)), vg.ISING, vg.FLIP, H, i)
             for i, L in enumerate(sides)]
    therm = [(float(T), abs(float(h)), (0.0, 0.0, 1.0 if h >= 0 else -1.0))
             for T, h in zip(temps, fields)]
    return sides, descs, therm


def run_batched(vg, ctx, sweeps, baseline_sweeps, n_chains=512, cpu=None):
    """All chains in one batch (one launch per colour pass for every chain,
    sweep blocks replayed from hipGraphs, tile autotuned) vs the same chains
    run one after another, each in its own batch with plain launches."""
    sides, descs, therm = batched_chains(vg, n_chains)
    n_sites = [int(L) ** 3 for L in sides]
    total = sum(n_sites)
    log(f"batched: {n_chains} chains, {total} sites, {sweeps} sweeps")
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        for i, (T, h, d) in enumerate(therm):
            b.set_thermostat(i, T, h, d)
        tile = b.autotune(30)  # 30 trial sweeps per tile: a stable pick
        b.set_tile(tile)
        b.set_graph(True)
        b.run(10)
        b.prepare(sweeps)
        t0 = time.perf_counter()
        b.run(sweeps)
        wall = time.perf_counter() - t0
    batched = total * sweeps / wall
    log(f"batched: {batched:.4e} attempts/s (tile {tile}, {wall:.2f} s)")
    # Per-chain launches: chain i alone, plain launches, timed run only.
    t_sum, att = 0.0, 0
    for i, d in enumerate(descs):
        with vg.Batch(ctx, [d]) as b1:
            b1.randomize()
            b1.set_thermostat(0, *therm[i])
            b1.run(2)
            t0 = time.perf_counter()
            b1.run(baseline_sweeps)
            t_sum += time.perf_counter() - t0
            att += n_sites[i] * baseline_sweeps
    per_chain = att / t_sum
    log(f"batched: per-chain plain launches {per_chain:.4e} attempts/s ({t_sum:.2f} s)")
    return {
        "workload": (f"{n_chains} Ising chains sc L^3, L uniform in {{16..128}} (numpy PCG64 "
                     f"seed 2024), T in [3.5,5.5], h in [-0.1,0.1], seeds 0..{n_chains - 1}"),
        "sites": total,
        "sweeps": sweeps,
        "tile": tile,
        "value": batched,
        "unit": "spin-flip attempts/s",
        "per_chain_launches": per_chain,
        "per_chain_sweeps": baseline_sweeps,
        "per_chain_sampling": (f"each chain alone for {baseline_sweeps} of the batch's {sweeps} "
                               f"sweeps (timed after 2 untimed ones): a chain's per-sweep cost "
                               f"does not depend on how many sweeps it runs"),
        "speedup": batched / per_chain,
        "cpu_all_cores": cpu,
    }


def run_bitpacked(vg, ctx, side, sweeps):
    """The headline chains (sc side^3 x 3, J = 1, h = 0, seed 7, T = 3.5,
    4.5115, 6.0) stored one bit per site; identical trajectories to int8."""
    temps = [3.5, 4.5115, 6.0]
    L = vg.lattice(vg.SC, side, side, side)
    descs = [vg.chain(L, vg.ISING_BITS, vg.FLIP, vg.hamiltonian(exchange=1.0), 7) for _ in temps]
    n = side ** 3
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        for i, T in enumerate(temps):
            b.set_thermostat(i, T)
        tile = b.autotune(50)
        b.set_tile(tile)
        b.set_graph(True)
        b.run(10)
        b.prepare(sweeps)
        t0 = time.perf_counter()
        b.run(sweeps)
        wall = time.perf_counter() - t0
        dev_ms = b.last_run_ms()
    value = n * len(temps) * sweeps / wall
    per_launch = n * len(temps) // 2
    avg_ms = dev_ms / (2 * sweeps)
    achieved = 0.25 * per_launch / (avg_ms * 1e-3) / 1e9
    log(f"bitpacked: {value:.4e} attempts/s, pass {avg_ms * 1e3:.1f} us, tile {tile}")
    return {
        "workload": f"ising-sc{side}-bits: the headline chains, 1 bit per site",
        "value": value,
        "unit": "spin-flip attempts/s",
        "sweeps": sweeps,
        "tile": tile,
        "roofline": with_counters(
            {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": achieved / HBM_PEAK_GBS, "kernel": "ising_sc_pass<bits>",
             "algorithmic_bytes_per_launch": per_launch // 4, "mean_launch_us": avg_ms * 1e3},
            "bitpacked", [f"ising_sc_pass<{tile},1>"], ising_mads(per_launch)),
    }


def run_resident(vg, ctx, side, sweeps):
    """One int8 Ising chain of sc side^3 at T_c = 4.5115 (J = 1, h = 0, seed
    7): at 1024^3 the two colour planes (1.07 GB) are far larger than the
    256 MB Infinity Cache, so every pass streams from HBM.  Attempts/s and
    the pass kernel's roofline at 2 B per attempt."""
    L = vg.lattice(vg.SC, side, side, side)
    n = side ** 3
    with vg.Batch(ctx, [vg.chain(L, vg.ISING, vg.FLIP, vg.hamiltonian(exchange=1.0), 7)]) as b:
        b.randomize()
        b.set_thermostat(0, 4.5115)
        tile = b.autotune(4)
        b.set_tile(tile)
        b.set_graph(True)
        b.run(4)
        b.prepare(sweeps)
        t0 = time.perf_counter()
        b.run(sweeps)
        wall = time.perf_counter() - t0
        dev_ms = b.last_run_ms()
    value = n * sweeps / wall
    per_launch = n // 2
    avg_ms = dev_ms / (2 * sweeps)
    achieved = 2 * per_launch / (avg_ms * 1e-3) / 1e9
    log(f"resident sc {side}^3: {value:.4e} attempts/s, pass {avg_ms * 1e3:.1f} us, tile {tile}")
    return {
        "workload": f"ising-sc{side}-i8: one chain sc {side}^3 at T=4.5115, HBM-resident",
        "value": value,
        "unit": "spin-flip attempts/s",
        "sweeps": sweeps,
        "tile": tile,
        "roofline": with_counters(
            {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": achieved / HBM_PEAK_GBS, "kernel": "ising_sc_pass",
             "algorithmic_bytes_per_launch": 2 * per_launch, "mean_launch_us": avg_ms * 1e3},
            "resident", [f"ising_sc_pass<{tile},2>"], ising_mads(per_launch)),
    }


def heis_mads(attempts_per_launch):
    """v_mad_u64_u32 (wave instructions) of a Heisenberg pass launch with
    uniform-sphere proposals: one Philox call (20) per site for the first two
    Marsaglia tries, 64 lanes per wave instruction (the 4.6% of sites whose
    tries both miss draw more, left out)."""
    return 20.0 * attempts_per_launch / 64.0


def run_heisenberg(vg, ctx, side, sweeps):
    """Heisenberg sc side^3 x 3 chains (J = 1, K = 0.1 along z, B = 0.05 along
    z, uniform-sphere proposals, random start from seed 7) at T = 0.5, 1.4,
    2.5; attempts/s and the pass kernel's roofline at 24 B per attempt."""
    temps = [0.5, 1.4, 2.5]
    H = vg.hamiltonian(exchange=1.0, zeeman=True, anisotropy=0.1)
    L = vg.lattice(vg.SC, side, side, side)
    descs = [vg.chain(L, vg.HEISENBERG, vg.RANDOM, H, 7) for _ in temps]
    n = side ** 3
    with vg.Batch(ctx, descs) as b:
        b.randomize()
        for i, T in enumerate(temps):
            b.set_thermostat(i, T, 0.05)
        b.set_graph(True)
        b.run(5)
        b.prepare(sweeps)
        t0 = time.perf_counter()
        e, m = b.run(sweeps)
        wall = time.perf_counter() - t0
        dev_ms = b.last_run_ms()
        b.set_graph(False)
        b.set_timing(True)
        b.run(20, observe=False)
        ev_ms, ev_n = b.pass_stats(vg.HEISENBERG)
    value = n * len(temps) * sweeps / wall
    per_launch = n * len(temps) // 2
    avg_ms = dev_ms / (2 * sweeps)
    achieved = 24 * per_launch / (avg_ms * 1e-3) / 1e9
    log(f"heisenberg: {value:.4e} attempts/s, pass {avg_ms * 1e3:.1f} us -> {achieved:.0f} GB/s")
    return {
        "workload": (f"heis-sc{side}-f32: 3 chains sc {side}^3, J=1, K=0.1 z, B=0.05 z, "
                     f"uniform-sphere proposals, T={temps}"),
        "value": value,
        "unit": "spin-flip attempts/s",
        "sweeps": sweeps,
        "ms_per_step": wall * 1e3 / sweeps,
        "roofline": with_counters(
            {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": achieved / HBM_PEAK_GBS, "kernel": "heis_sc_pass",
             "algorithmic_bytes_per_launch": 24 * per_launch, "mean_launch_us": avg_ms * 1e3,
             "mean_launch_us_per_launch_events": ev_ms / ev_n * 1e3},
            "heisenberg", ["heis_sc_pass"], heis_mads(per_launch)),
        "observables_last_sweep": {f"T={T}": {"E/N": float(e[i, -1]) / n,
                                              "|M|/N": float(m[i, -1]) / n}
                                   for i, T in enumerate(temps)},
    }


def pass_kernels(lattice, spin, tile):
    """The colour-pass kernels of a bcc or fcc line as the counter table names
    them (fcc: one instantiation per colour)."""
    bits = 1 if spin == "ising_bits" else 0
    if lattice == "bcc":
        return ["heis_bcc_pass"] if spin == "heisenberg" else [f"ising_bcc_pass<{tile},{bits}>"]
    if spin == "heisenberg":
        return [f"heis_fcc_pass<{a}>" for a in range(4)]
    return [f"ising_fcc_pass<{a},{tile},{bits}>" for a in range(4)]


def run_fcc(vg, ctx, cells, heis_cells, sweeps, kind=None):
    """fcc lattices (four colour passes per sweep): 3 Ising chains of cells^3
    fcc cells (4 sites each; J = 1, h = 0; T = 8, 9.8, 12 around T_c ~ 9.8)
    and 3 Heisenberg chains of heis_cells^3 cells (J = 1, K = 0.1 z, B = 0.05
    z, uniform-sphere proposals; T = 1.5, 3.2, 5), random starts from seed 7.
    The Ising chains run again stored one bit per site.  Attempts/s and each
    pass kernel's roofline (2 B / 0.25 B / 24 B per attempt).
    With kind = vg.BCC the same on bcc lattices (2 sites per cell, two colour
    passes; Ising T = 5, 6.35, 8 around T_c ~ 6.35, Heisenberg T = 1, 2.1, 3.5)."""
    bcc = kind == vg.BCC
    kind = vg.BCC if bcc else vg.FCC
    per_cell, passes, name = (2, 2, "bcc") if bcc else (4, 4, "fcc")
    out = {}
    spins = [("ising", cells, [5.0, 6.35, 8.0] if bcc else [8.0, 9.8, 12.0]),
             ("heisenberg", heis_cells, [1.0, 2.1, 3.5] if bcc else [1.5, 3.2, 5.0])]
    # the Ising chains again, stored one bit per site
    spins.append(("ising_bits", cells, spins[0][2]))
    for spin, side, temps in spins:
        ising = spin != "heisenberg"
        H = (vg.hamiltonian(exchange=1.0) if ising
             else vg.hamiltonian(exchange=1.0, zeeman=True, anisotropy=0.1))
        L = vg.lattice(kind, side, side, side)
        kind_spin = {"ising": vg.ISING, "ising_bits": vg.ISING_BITS, "heisenberg": vg.HEISENBERG}
        descs = [vg.chain(L, kind_spin[spin], vg.FLIP if ising else vg.RANDOM, H, 7)
                 for _ in temps]
        n = per_cell * side ** 3
        with vg.Batch(ctx, descs) as b:
            b.randomize()
            for i, T in enumerate(temps):
                b.set_thermostat(i, T, 0.0 if ising else 0.05)
            tile = b.autotune(30)  # the Ising passes' z tile (a no-op for Heisenberg)
            b.set_graph(True)
            b.run(5)
            b.prepare(sweeps)
            t0 = time.perf_counter()
            e, m = b.run(sweeps)
            wall = time.perf_counter() - t0
            dev_ms = b.last_run_ms()
        value = n * len(temps) * sweeps / wall
        per_launch = n * len(temps) // passes
        avg_ms = dev_ms / (passes * sweeps)
        bpa = {"ising": 2, "ising_bits": 0.25, "heisenberg": 24}[spin]
        achieved = bpa * per_launch / (avg_ms * 1e-3) / 1e9
        log(f"{name} {spin}: {value:.4e} attempts/s, pass {avg_ms * 1e3:.1f} us -> "
            f"{achieved:.0f} GB/s")
        out[spin] = {
            "workload": (f"{spin}-{name}{side}: 3 chains of {side}^3 {name} cells ({n} sites), "
                         f"T={temps}" + ("" if ising else ", K=0.1 z, B=0.05 z, random proposals")),
            "value": value,
            "unit": "spin-flip attempts/s",
            "sweeps": sweeps,
            "ms_per_step": wall * 1e3 / sweeps,
            "roofline": with_counters(
                {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                 "kernel": (f"{'ising' if ising else 'heis'}_{name}_pass"
                            + ("<bits>" if spin == "ising_bits" else "")),
                 "algorithmic_bytes_per_launch": bpa * per_launch,
                 "mean_launch_us": avg_ms * 1e3},
                name, pass_kernels(name, spin, tile),
                ising_mads(per_launch) if ising else heis_mads(per_launch)),
            "observables_last_sweep": {f"T={T}": {"E/N": float(e[i, -1]) / n,
                                                  "|M|/N": float(m[i, -1]) / n}
                                       for i, T in enumerate(temps)},
        }
    return out


def run_wolff(vg, ctx, batch_side, batch_chains, single_side, steps, cpu_seconds=5.0):
    """Wolff cluster steps (integrator.rs:146-230) at T_c = 4.5115 of the sc
    Ising model, J = 1: batch_chains chains of batch_side^3 (one workgroup
    per chain) and one chain of single_side^3, random starts, 1000 steps of
    warm-up (the single large chain is still far from equilibrium then, so
    its clusters are small: the figure is the per-step cost).  Reports cluster steps/s and flipped sites/s (a cluster's size
    is half the change of the signed magnetization), and the oracle's Wolff
    (the same clusters) on one core for the single chain as the CPU figure."""
    import numpy as np
    T = 4.5115
    out = {}
    for name, side, nch in (("batched", batch_side, batch_chains), ("single", single_side, 1)):
        L = vg.lattice(vg.SC, side, side, side)
        descs = [vg.chain(L, vg.ISING, vg.WOLFF, vg.hamiltonian(exchange=1.0), s)
                 for s in range(nch)]
        with vg.Batch(ctx, descs) as b:
            b.randomize()
            for i in range(nch):
                b.set_thermostat(i, T)
            b.set_graph(True)
            b.run(1000, observe=False)
            m0 = b.run_sums(1)[:, -1, 3]
            b.prepare(steps)
            t0 = time.perf_counter()
            sums = b.run_sums(steps)
            wall = time.perf_counter() - t0
        mags = np.concatenate([m0[:, None], sums[:, :, 3]], axis=1)
        flipped = int(np.abs(np.diff(mags, axis=1)).sum() // 2)
        out[name] = {
            "workload": f"{nch} chains sc {side}^3 at T={T}, {steps} Wolff steps each",
            "steps_per_s": nch * steps / wall,
            "flipped_sites_per_s": flipped / wall,
            "mean_cluster_size": flipped / (nch * steps),
            "ms_per_step": wall * 1e3 / steps,
        }
        log(f"wolff {name}: {nch * steps / wall:.3e} steps/s, {flipped / wall:.3e} flipped sites/s, "
            f"mean cluster {flipped / (nch * steps):.1f}")
    # CPU: the oracle's Wolff on the single chain, for about cpu_seconds
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    OL = orc.lattice("sc", single_side, single_side, single_side)
    OM = orc.model(terms=("exchange",), temperature=T)
    sp = orc.ising_init(OL, 0)
    orc.wolff_steps(OL, OM, 0, 0, 200, sp)
    done, flipped, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < cpu_seconds:
        m = orc.wolff_steps(OL, OM, 0, 200 + done, 50, sp)[:, 3]
        flipped += int(np.abs(np.diff(m)).sum() // 2)
        done += 50
    dt = time.perf_counter() - t0
    out["cpu"] = {"steps_per_s": done / dt, "flipped_sites_per_s": flipped / dt, "cores": 1,
                  "kind": "port",
                  "sample": (f"{done} Wolff steps of sc {single_side}^3 at T={T} (oracle "
                             f"restatement of WolffIntegrator::step, 1 thread on {cpu_model()})")}
    log(f"wolff cpu: {done / dt:.3e} steps/s, {flipped / dt:.3e} flipped sites/s")
    return out


def served_jobs(vg, n_jobs, gen_seed=2025):
    """The served mix: 60% Ising Relax (sc L^3, L even in [16, 96], T uniform in
    [2, 6], 2000 sweeps), 30% Heisenberg CoolDown (sc L^3, L even in [16, 64],
    K = 0.1 along z, T 3.0 -> 0.1 by 0.1 = 29 temperatures, 500 measured
    sweeps each), 10% Heisenberg HysteresisLoop (sc 64^3, K = 0.1, T = 1.0,
    field 0 -> 0.5 -> -0.5 -> 0.5 by 0.05 = 55 stages, 1000 measured sweeps
    each).  Compound(Exchange J=1, Zeeman[, Anisotropy]); seeds = job index.
    Even L: a periodic sc axis needs even length to be two-colourable."""
    import numpy as np
    rng = np.random.default_rng(gen_seed)
    ising_h = vg.hamiltonian(exchange=1.0, zeeman=True)
    heis_h = vg.hamiltonian(exchange=1.0, zeeman=True, anisotropy=0.1)
    cool = vg.cool_down(3.0, 0.1, 0.1, relax=0, steps=500)
    loop = vg.hysteresis_loop(steps=1000, relax=0, temperature=1.0, max_field=0.5,
                              field_step=0.05)
    jobs = []
    for k in range(n_jobs):
        u = rng.random()
        if u < 0.6:
            L = int(rng.integers(8, 49)) * 2
            T = float(rng.uniform(2.0, 6.0))
            d = vg.chain(vg.lattice(vg.SC, L, L, L), vg.ISING, vg.FLIP, ising_h, k)
            jobs.append(("ising-relax", d, vg.relax(2000, T), L ** 3 * 2000, ("ising", L, 2000)))
        elif u < 0.9:
            L = int(rng.integers(8, 33)) * 2
            d = vg.chain(vg.lattice(vg.SC, L, L, L), vg.HEISENBERG, vg.RANDOM, heis_h, k)
            jobs.append(("heis-cooldown", d, cool, L ** 3 * 500 * len(cool),
                         ("heis", L, 500 * len(cool))))
        else:
            d = vg.chain(vg.lattice(vg.SC, 64, 64, 64), vg.HEISENBERG, vg.RANDOM, heis_h, k)
            jobs.append(("heis-hysteresis", d, loop, 64 ** 3 * 1000 * len(loop),
                         ("heis", 64, 1000 * len(loop))))
    return jobs


def one_at_a_time_seconds(vg, ctx, jobs, probe_sweeps=60):
    """Device time the job list would take run one job after another, each
    as its own one-chain batch with plain launches: per-sweep time measured
    for every (spin kind, L) in the list over probe_sweeps sweeps, times each
    job's sweeps."""
    per_sweep = {}
    ising_h = vg.hamiltonian(exchange=1.0, zeeman=True)
    heis_h = vg.hamiltonian(exchange=1.0, zeeman=True, anisotropy=0.1)
    for key in sorted({j[4][:2] for j in jobs}):
        spin, L = key
        lat = vg.lattice(vg.SC, L, L, L)
        d = (vg.chain(lat, vg.ISING, vg.FLIP, ising_h, 1) if spin == "ising" else
             vg.chain(lat, vg.HEISENBERG, vg.RANDOM, heis_h, 1))
        with vg.Batch(ctx, [d]) as b:
            b.randomize()
            b.set_thermostat(0, 2.0, 0.05)
            b.run(5, observe=False)
            t0 = time.perf_counter()
            b.run(probe_sweeps)
            per_sweep[key] = (time.perf_counter() - t0) / probe_sweeps
    return sum(j[4][2] * per_sweep[j[4][:2]] for j in jobs)


def run_served(vg, ctx, n_jobs, rates, gen_seed=2025, budget=0, graph=False):
    """Poisson arrivals at each rate: one thread submits each job at its
    arrival time; latency = finish - arrival.  Throughput = attempts of all
    jobs / (last finish - first arrival).  graph: the server replays quanta
    from hipGraphs while the running set is unchanged."""
    import numpy as np
    jobs = served_jobs(vg, n_jobs, gen_seed)
    out = []
    for rate in rates:
        arr_rng = np.random.default_rng(gen_seed + int(rate))
        arrivals = np.cumsum(arr_rng.exponential(1.0 / rate, size=len(jobs)))
        log(f"served: {len(jobs)} jobs at {rate}/s")
        with vg.Server(ctx, memory_bytes=16 << 30, max_jobs=1024, quantum_sweeps=100,
                       tile=2, use_graph=graph, sweep_budget=budget) as srv:
            t0 = time.perf_counter()
            ids, sub_t = [], []
            for (kind, d, stages, _, _), ta in zip(jobs, arrivals):
                delay = t0 + ta - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids.append(srv.submit(d, stages))
                sub_t.append(time.perf_counter() - t0)
            ends = []
            for jid in ids:
                st = srv.wait(jid, 1200.0)
                if st.state != vg.JOB_DONE:
                    raise RuntimeError(f"job {jid} ended in state {st.state}")
                ends.append(st.end_s)
            stats = srv.stats()
            # server clock: seconds since its start, which precedes t0 slightly
            first = srv.status(ids[0])
            offset = first.submit_s - sub_t[0]
            lat = np.array([e - (ta + offset) for e, ta in zip(ends, arrivals)])
            makespan = max(ends) - (arrivals[0] + offset)
        attempts = sum(j[3] for j in jobs)
        kinds = {}
        for (kind, *_), l in zip(jobs, lat):
            kinds.setdefault(kind, []).append(l)
        res = {
            "rate_per_s": rate,
            "jobs": len(jobs),
            "sweep_budget": budget,
            "attempts": attempts,
            "throughput": attempts / makespan,
            "makespan_s": makespan,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "p50_by_kind_s": {k: float(np.percentile(v, 50)) for k, v in kinds.items()},
            "p99_by_kind_s": {k: float(np.percentile(v, 99)) for k, v in kinds.items()},
            "peak_jobs": stats.peak_jobs,
            "pool_peak_bytes": stats.pool_peak,
            "device_busy_frac": stats.device_ms / 1e3 / makespan,
        }
        log(f"served: {rate}/s -> {res['throughput']:.3e} attempts/s, p50 "
            f"{res['p50_latency_s']:.3f} s, p99 {res['p99_latency_s']:.3f} s, peak "
            f"{stats.peak_jobs} jobs")
        out.append(res)
    # The same jobs one at a time with per-chain plain launches.
    seq_s = one_at_a_time_seconds(vg, ctx, jobs)
    attempts = sum(j[3] for j in jobs)
    for res in out:
        res["one_at_a_time_throughput_estimated"] = attempts / seq_s
        res["speedup_vs_one_at_a_time_estimated"] = res["throughput"] / (attempts / seq_s)
        res["one_at_a_time_method"] = (
            "estimated, not run end to end: the per-sweep wall time of a one-chain batch with "
            "plain launches, measured over 60 sweeps for every (spin kind, L) in the job list, "
            "times each job's sweeps, summed over the jobs")
    log(f"served: one job at a time with plain launches, estimated {attempts / seq_s:.3e} "
        f"attempts/s ({seq_s:.1f} s)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--side", type=int, default=256, help="lattice side (256 for the headline)")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="plain launches instead of replaying sweep blocks from a hipGraph")
    ap.add_argument("--cpu-sweeps", type=int, default=5,
                    help="timed sweeps of the CPU baseline (after one warm-up sweep)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-sweeps", type=int, default=100)
    ap.add_argument("--tile", type=int, default=0, help="Ising tile; 0 = autotune")
    ap.add_argument("--workload", default="all",
                    choices=["all", "headline", "batched", "served", "heisenberg", "bitpacked",
                             "fcc", "bcc", "wolff", "resident"],
                    help="headline: sc 256^3 x 3 only; batched: the 512-chain batch only; "
                         "served: the Poisson job stream only")
    ap.add_argument("--serve-jobs", type=int, default=2000, help="jobs per rate")
    ap.add_argument("--heis-side", type=int, default=128)
    ap.add_argument("--heis-sweeps", type=int, default=500)
    ap.add_argument("--fcc-cells", type=int, default=128, help="fcc Ising chains: cells per side")
    ap.add_argument("--fcc-heis-cells", type=int, default=80,
                    help="fcc Heisenberg chains: cells per side")
    ap.add_argument("--fcc-sweeps", type=int, default=300, help="sweeps of the fcc and bcc lines")
    ap.add_argument("--bcc-cells", type=int, default=160, help="bcc Ising chains: cells per side")
    ap.add_argument("--bcc-heis-cells", type=int, default=100,
                    help="bcc Heisenberg chains: cells per side")
    ap.add_argument("--wolff-steps", type=int, default=2000)
    ap.add_argument("--resident-side", type=int, default=1024,
                    help="side of the HBM-resident Ising chain")
    ap.add_argument("--resident-sweeps", type=int, default=100)
    ap.add_argument("--serve-rates", default="50,100,200,500")
    ap.add_argument("--serve-graph", action="store_true",
                    help="the server replays quanta from hipGraphs")
    ap.add_argument("--serve-budget", type=int, default=0,
                    help="server per-sweep work budget (Ising-site units; 0 = none)")
    ap.add_argument("--batched-sweeps", type=int, default=10000)
    ap.add_argument("--baseline-sweeps", type=int, default=2000,
                    help="sweeps of each chain in the batched set's per-chain leg")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.copy_bytes:
        import torch
        a = torch.ones(args.copy_bytes // 4, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); every chain "
              f"fits one MI355X and multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    # The batched set's all-cores CPU aggregate forks worker processes, so it
    # runs before anything touches the GPU.
    cpu_batched = None
    if not args.no_cpu and args.workload in ("all", "batched"):
        log(f"bench: CPU aggregate of the batched set ({cpu_threads()} processes)")
        cpu_batched = cpu_batched_aggregate(cpu_threads())
    vg = load_engine()
    if args.workload == "batched":
        with vg.Context(0) as ctx:
            print(json.dumps(run_batched(vg, ctx, args.batched_sweeps, args.baseline_sweeps,
                                         cpu=cpu_batched)), flush=True)
        return
    if args.workload == "heisenberg":
        with vg.Context(0) as ctx:
            print(json.dumps(run_heisenberg(vg, ctx, args.heis_side, args.heis_sweeps)),
                  flush=True)
        return
    if args.workload == "wolff":
        with vg.Context(0) as ctx:
            print(json.dumps(run_wolff(vg, ctx, 32, 64, 64, args.wolff_steps)), flush=True)
        return
    if args.workload == "resident":
        with vg.Context(0) as ctx:
            print(json.dumps(run_resident(vg, ctx, args.resident_side, args.resident_sweeps)),
                  flush=True)
        return
    if args.workload == "fcc":
        with vg.Context(0) as ctx:
            print(json.dumps(run_fcc(vg, ctx, args.fcc_cells, args.fcc_heis_cells,
                                     args.fcc_sweeps)), flush=True)
        return
    if args.workload == "bcc":
        with vg.Context(0) as ctx:
            print(json.dumps(run_fcc(vg, ctx, args.bcc_cells, args.bcc_heis_cells,
                                     args.fcc_sweeps, kind=vg.BCC)), flush=True)
        return
    if args.workload == "bitpacked":
        with vg.Context(0) as ctx:
            print(json.dumps(run_bitpacked(vg, ctx, args.side, args.steps)), flush=True)
        return
    if args.workload == "served":
        rates = [float(r) for r in args.serve_rates.split(",")]
        with vg.Context(0) as ctx:
            print(json.dumps(run_served(vg, ctx, args.serve_jobs, rates, graph=args.serve_graph,
                                        budget=args.serve_budget)), flush=True)
        return
    temps = [3.5, 4.5115, 6.0]
    side = args.side
    L = vg.lattice(vg.SC, side, side, side)
    H = vg.hamiltonian(exchange=1.0)
    descs = [vg.chain(L, vg.ISING, vg.FLIP, H, 7) for _ in temps]
    n_sites = side ** 3
    attempts_per_step = n_sites * len(temps)
    ctx = vg.Context(0)
    batch = vg.Batch(ctx, descs)
    batch.randomize()
    for i, T in enumerate(temps):
        batch.set_thermostat(i, T)
    tile = batch.autotune(50) if args.tile == 0 else args.tile
    batch.set_tile(tile)
    batch.set_graph(args.graph)
    log(f"bench: {len(temps)} chains sc {side}^3 int8, warmup {args.warmup}")
    if args.warmup:
        batch.run(args.warmup)
    # The sweep-block graphs of the timed run are captured and uploaded here,
    # outside the timed region (no sweep runs).
    batch.prepare(args.steps)
    log(f"bench: timing {args.steps} sweeps")
    t0 = time.perf_counter()
    e, m = batch.run(args.steps)
    wall = time.perf_counter() - t0
    device_ms = batch.last_run_ms()
    value = attempts_per_step * args.steps / wall
    log(f"bench: {value:.4e} attempts/s (wall {wall:.2f} s, device {device_ms / 1e3:.2f} s)")

    # Mean launch time of the colour pass from the HIP events the engine
    # records on its stream around every timed sweep block: the blocks hold
    # 2 pass launches per sweep plus one small reduction, so this slightly
    # overstates the pass time.
    launches = 2 * args.steps
    avg_ms = device_ms / launches
    # Cross-check: per-launch events around each pass of a short plain run.
    batch.set_graph(False)
    batch.set_timing(True)
    tsw = max(1, min(args.timing_sweeps, args.steps))
    batch.run(tsw, observe=False)
    ev_ms, ev_launches = batch.pass_stats(vg.ISING)
    batch.set_timing(False)
    attempts_per_launch = attempts_per_step // 2
    alg_bytes = 2 * attempts_per_launch  # read + write of each site once per sweep
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    log(f"bench: pass kernel {avg_ms * 1e3:.1f} us/launch -> {achieved:.0f} GB/s algorithmic")

    cpu = None
    if not args.no_cpu:
        log(f"bench: CPU baseline ({args.cpu_sweeps} sweeps at sc {side}^3)")
        cpu = cpu_baseline(side, 4.5115, args.cpu_sweeps)
        log("bench: best-CPU line (the oracle's fast checkerboard, 1 core and all cores)")
        cpu["best_cpu_checkerboard"] = cpu_best_line(side, 4.5115, 2, 10, cpu_threads())

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "spin-flip attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic (Philox random +-1 start, seed 7)",
        "config": {
            "workload": (f"ising-sc{side}-i8: {len(temps)} chains sc {side}x{side}x{side}, full "
                         f"PBC, J=1, h=0, int8 spins, seed 7, T={temps}, checkerboard flip "
                         f"sweeps; one step = one sweep of every chain"),
            "lattice": f"sc {side}^3",
            "chains": len(temps),
            "attempts_per_step": attempts_per_step,
            "graph": bool(args.graph),
            "tile": tile,
        },
        "roofline": with_counters({
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "kernel": "ising_sc_pass",
            "algorithmic_bytes_per_launch": alg_bytes,
            "mean_launch_us": avg_ms * 1e3,
            "launches_timed": launches,
            "mean_launch_us_per_launch_events": ev_ms / ev_launches * 1e3,
        }, "headline" if side == 256 else f"sc{side}", [f"ising_sc_pass<{tile},2>"],
            ising_mads(attempts_per_launch)),
        "cpu_baseline": cpu,
        "observables_last_sweep": {
            f"T={T}": {"E/N": float(e[i, -1]) / n_sites, "|M|/N": float(m[i, -1]) / n_sites}
            for i, T in enumerate(temps)
        },
    }
    batch.close()
    if args.workload == "all":
        # The other forms of the path, each its own sub-object (not `value`).
        out["bitpacked"] = run_bitpacked(vg, ctx, side, args.steps)
        out["resident"] = run_resident(vg, ctx, args.resident_side, args.resident_sweeps)
        out["batched"] = run_batched(vg, ctx, args.batched_sweeps, args.baseline_sweeps,
                                     cpu=cpu_batched)
        out["heisenberg"] = run_heisenberg(vg, ctx, args.heis_side, args.heis_sweeps)
        out["fcc"] = run_fcc(vg, ctx, args.fcc_cells, args.fcc_heis_cells, args.fcc_sweeps)
        out["bcc"] = run_fcc(vg, ctx, args.bcc_cells, args.bcc_heis_cells, args.fcc_sweeps,
                             kind=vg.BCC)
        out["wolff"] = run_wolff(vg, ctx, 32, 64, 64, args.wolff_steps)
        rates = [float(r) for r in args.serve_rates.split(",")]
        out["served"] = run_served(vg, ctx, args.serve_jobs, rates, budget=args.serve_budget,
                                   graph=args.serve_graph)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
