"""Exact thermal averages of the 10x10 periodic Ising model -- the reference's
CPU-runnable configuration (sc 10x10x1, x and y periodic, z open, J = 1,
h = 0) -- at the temperatures of its program (Relax at 4.0, CoolDown
4.0 -> 0.1 by 0.1: 39 stages, the f64 schedule of src/program.rs:202-211).

The dynamics sample exp(-H/T) with H = -sum_bonds s_i s_j; the observable
energy is the Compound(Exchange, Zeeman) total, sum_i e_i = 2 H
(src/energy.rs:55-59), so E/N = 2 <H>/N.  Row-to-row transfer matrix over
the 2^10 configurations of a row: W[a, b] = exp(beta (bonds(a, b) +
(bonds(a) + bonds(b)) / 2)), Z = Tr W^10.

- <E>/N and C = Var(E)/(N T^2): from ln Z(beta) = ln sum_i lambda_i^10
  (eigenvalues of the symmetric W) by central differences in beta
  (Var(E_obs) = 4 Var(H) = 4 d^2 ln Z / d beta^2).
- <|M|>/N and chi = Var(|M|)/(N T): from the distribution of M = sum s,
  P(M) = (1/101) sum_k Z(theta_k) e^{-i theta_k M} / Z(0) over the 101
  values M = -100, -98, ..., 100, where Z(theta) = Tr (D W D)^10 with
  D = diag(e^{i theta m(a) / 2}) and theta_k = 2 pi k / 202.

Writes ising10x10_exact.json beside this script.  Run: python
make_ising10x10_exact.py [workers]."""
import itertools
import json
import math
import sys
from multiprocessing import Pool
from pathlib import Path

import numpy as np

L = 10
CFG = np.array(list(itertools.product([-1.0, 1.0], repeat=L)))
MROW = CFG.sum(axis=1)
INTRA = (CFG * np.roll(CFG, 1, axis=1)).sum(axis=1)
INTER = CFG @ CFG.T
BONDS = INTER + 0.5 * (INTRA[:, None] + INTRA[None, :])


def ln_z(beta):
    """ln Tr W^L with W scaled by its largest eigenvalue for range."""
    lam = np.linalg.eigvalsh(np.exp(beta * BONDS))
    top = lam.max()
    return L * math.log(top) + math.log(float(np.sum((lam / top) ** L)))


def m_distribution(beta):
    W = np.exp(beta * BONDS)
    W /= np.linalg.eigvalsh(W).max()
    K = 101
    ms = np.arange(-L * L, L * L + 1, 2)
    zs = []
    for k in range(K):
        theta = 2.0 * math.pi * k / (2 * K)
        d = np.exp(0.5j * theta * MROW)
        A = W * np.outer(d, d)
        A2 = A @ A
        A4 = A2 @ A2
        A8 = A4 @ A4
        zs.append(np.trace(A8 @ A2))
    zs = np.array(zs)
    thetas = 2.0 * math.pi * np.arange(K) / (2 * K)
    p = np.array([np.sum(zs * np.exp(-1j * thetas * m)).real for m in ms]) / K
    p /= p.sum()
    return ms, np.clip(p, 0.0, None)


def exact(T):
    beta = 1.0 / T
    eps = 1e-4 * beta
    lp, l0, lm = ln_z(beta + eps), ln_z(beta), ln_z(beta - eps)
    h_mean = -(lp - lm) / (2 * eps)              # <H> = -d ln Z / d beta
    h_var = (lp - 2 * l0 + lm) / (eps * eps)     # Var(H) = d^2 ln Z / d beta^2
    n = L * L
    ms, p = m_distribution(beta)
    am = float(np.sum(np.abs(ms) * p))
    m2 = float(np.sum(ms * ms * p))
    return {"T": T, "E/N": 2.0 * h_mean / n, "|M|/N": am / n,
            "C": 4.0 * h_var / (n * T * T), "chi": (m2 - am * am) / (n * T)}


def schedule():
    """Relax at 4.0, then CoolDown(4.0, 0.1, 0.1): t -= 0.1 after each stage
    until t < 0.1 (src/program.rs:202-211)."""
    temps, t = [], 4.0
    while True:
        temps.append(t)
        t -= 0.1
        if t < 0.1:
            break
    return temps


def main():
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 4
    temps = schedule()
    assert len(temps) == 39
    with Pool(workers) as pool:
        rows = pool.map(exact, temps)
    out = Path(__file__).resolve().parent / "ising10x10_exact.json"
    out.write_text(json.dumps({"lattice": "sc 10x10x1, x and y periodic, z open, J = 1, h = 0",
                               "method": "row transfer matrix (make_ising10x10_exact.py)",
                               "rows": rows}, indent=1))
    print(f"wrote {out} ({len(rows)} temperatures)")


if __name__ == "__main__":
    main()
