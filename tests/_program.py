"""A program (the stage list of Relax, CoolDown or HysteresisLoop) run on the
CPU oracle's checkerboard, with the reference's observables and StatSensor
accumulation -- the checker for the job server's per-stage results.

Per sweep, E and |M| follow Hamiltonian::total_energy and
State::magnetization (src/energy.rs:55-59, 114-120, 153-160, 208-213;
src/state.rs:86-92, 150-160, 291-296) evaluated on the oracle's exact
per-sweep sums; per measured stage the accumulators and the line are
StatSensor's (src/instrument.rs:93-123, src/accumulator.rs:42-63):
T, |h|, <E>, Var(E)/(N T^2), <|M|>, Var(|M|)/(N T), Binder(|M|).

Every floating-point step is done in the order written so that a correct
engine reproduces the values bit for bit from the same integer sums."""
import math

import numpy as np

EPS = float(np.finfo(float).eps)
HEIS_SCALE = 1.0 / 134217728.0  # 2^-27: fixed point of the Heisenberg sums
KINDS = {0: "sc", 1: "bcc", 2: "fcc"}
TERMS = ((1, "exchange"), (2, "zeeman"), (4, "anisotropy"), (8, "gauge"))


def oracle_lattice(orc, L):
    return orc.lattice(KINDS[L.kind], L.lx, L.ly, L.lz,
                       pbc=(bool(L.pbc & 1), bool(L.pbc & 2), bool(L.pbc & 4)))


def oracle_model(orc, d, T, field, field_dir):
    H = d.hamiltonian
    terms = tuple(name for bit, name in TERMS if H.terms & bit)
    M = orc.model(spin="heisenberg" if d.spin == 2 else "ising",
                  proposal="random" if d.proposal == 1 else "flip", terms=terms,
                  exchange=H.exchange, anisotropy=H.anisotropy,
                  axis=tuple(H.anisotropy_axis), gauge=H.gauge,
                  temperature=max(T, EPS), field=field, field_dir=tuple(field_dir))
    M.compound = H.compound
    return M


def observables(d, n_sites, sums, field, field_dir):
    """E and |M| of one sweep from its integer sums {bond, mx, my, mz, axis2}."""
    H = d.hamiltonian
    heis = d.spin == 2
    sc = HEIS_SCALE if heis else 1.0
    N = float(n_sites)
    bond = float(int(sums[0])) * sc
    m = [float(int(sums[1])) * sc, float(int(sums[2])) * sc, float(int(sums[3])) * sc]
    if heis:
        axis2 = float(int(sums[4])) * sc
        hdir = [float(field_dir[0]), float(field_dir[1]), float(field_dir[2])]
    else:
        axis2 = N
        hdir = [0.0, 0.0, 1.0 if field_dir[2] >= 0.0 else -1.0]
    hm = abs(field)
    mdoth = m[0] * hdir[0] + m[1] * hdir[1] + m[2] * hdir[2]
    e = 0.0
    if H.compound:
        if H.terms & 1:
            e += -2.0 * H.exchange * bond
        if H.terms & 2:
            e += hm * mdoth
        if H.terms & 4:
            e += H.anisotropy * axis2
        if H.terms & 8:
            e += H.gauge * N
    elif H.terms & 1:
        e = -H.exchange * bond
    elif H.terms & 2:
        e = -hm * mdoth
    elif H.terms & 4:
        e = axis2
    else:
        e = H.gauge * N
    if heis:
        mag = math.sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2])
        if abs(mag) < EPS:
            mag = 0.0
    else:
        mag = abs(m[2])
    return e, mag


def stat_line(T, field, n_sites, es, ms):
    """StatSensor's per-stage line from per-sweep E and |M| (sweep order)."""
    se = se2 = sm = sm2 = sm4 = 0.0
    for x, y in zip(es, ms):
        se += x
        se2 += x * x
        sm += y
        sm2 += y * y
        sm4 += y * y * y * y
    c = float(len(es))
    n = float(n_sites)
    if not len(es):
        return [T, abs(field), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    me, mm = se / c, sm / c
    m2 = sm2 / c
    return [T, abs(field), me, (se2 / c - me * me) / (n * (T * T)), mm,
            (sm2 / c - mm * mm) / (n * T), 1.0 - (sm4 / c) / (3.0 * (m2 * m2)), c]


def oracle_program(orc, d, stages):
    """Per-stage results of a job (chain descriptor d, vg stage list) on the
    oracle: initial state from the chain's seed, one sweep counter across
    the stages; rows as the server's results().  Also returns E and |M|
    after the job's last sweep (the server's last_energy and
    last_magnetization)."""
    OL = oracle_lattice(orc, d.lattice)
    n = orc.num_sites(OL)
    heis = d.spin == 2
    state = orc.heis_init(OL, d.seed) if heis else orc.ising_init(OL, d.seed)
    run = orc.heis_sweeps if heis else orc.ising_sweeps
    sweep = 0
    rows = []
    last = None
    for s in stages:
        T = max(s.temperature, EPS)
        fdir = tuple(s.field_dir)
        M = oracle_model(orc, d, T, s.field, fdir)
        es, ms = [], []
        if s.relax:
            sums = run(OL, M, d.seed, sweep, s.relax, state)
            sweep += s.relax
            last = observables(d, n, sums[-1], s.field, fdir)
        if s.measure:
            sums = run(OL, M, d.seed, sweep, s.measure, state)
            sweep += s.measure
            for row in sums:
                e, m = observables(d, n, row, s.field, fdir)
                es.append(e)
                ms.append(m)
            last = (es[-1], ms[-1])
        rows.append(stat_line(T, s.field, n, es, ms))
    return np.array(rows), last
