"""Binned error bars and exact enumeration used by the statistical tests."""
import itertools

import numpy as np


def binned(x, nbins=40):
    """Mean and standard error of a correlated series from `nbins` bin means."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x) // nbins * nbins
    bins = x[:n].reshape(nbins, -1).mean(axis=1)
    return float(bins.mean()), float(bins.std(ddof=1) / np.sqrt(nbins))


def binned_fn(f, series, nbins=40):
    """Error of a function of bin means (jackknife over bins)."""
    series = [np.asarray(s, dtype=np.float64) for s in series]
    n = len(series[0]) // nbins * nbins
    cut = [s[:n].reshape(nbins, -1) for s in series]
    full = f(*[c.reshape(-1) for c in cut])
    jack = []
    for k in range(nbins):
        keep = [np.delete(c, k, axis=0).reshape(-1) for c in cut]
        jack.append(f(*keep))
    jack = np.array(jack)
    err = np.sqrt((nbins - 1) / nbins * ((jack - jack.mean()) ** 2).sum())
    return float(full), float(err)


def z_score(a, ea, b, eb):
    return abs(a - b) / np.sqrt(ea ** 2 + eb ** 2 + 1e-300)


def ising_exact(neigh, n, J, h, hz, T, compound=True):
    """Exact Boltzmann averages over all 2^n Ising states.

    neigh[i] lists the neighbour slots of site i (with multiplicity); the
    per-site energy is e_i = -J s_i sum_j s_j + |h| hz s_i (the reference's
    Exchange + Zeeman, energy.rs:147-151,194-206) and dynamics sample
    exp(-sum_i e_i^dyn / T) where the Metropolis energy is the Hamiltonian
    H = -J sum_bonds s_i s_j + |h| hz sum_i s_i.  Returns <E_obs>, <|M|>,
    Var(E_obs), Var(|M|) with E_obs the reference's Compound total energy
    sum_i e_i (exchange counted twice)."""
    states = np.array(list(itertools.product([-1, 1], repeat=n)), dtype=np.float64)
    nb = np.zeros((len(states), n))
    for i in range(n):
        for j in neigh[i]:
            nb[:, i] += states[:, j]
    bond2 = (states * nb).sum(axis=1)          # sum_i s_i sum_j s_j = 2 * bonds
    m = states.sum(axis=1)
    H = -J * bond2 / 2 + abs(h) * hz * m       # energy the dynamics sample
    if compound:
        e_obs = -J * bond2 + abs(h) * hz * m
    else:
        e_obs = -J * bond2 / 2
    w = np.exp(-(H - H.min()) / T)
    Z = w.sum()
    mean_e = (w * e_obs).sum() / Z
    mean_m = (w * np.abs(m)).sum() / Z
    var_e = (w * e_obs ** 2).sum() / Z - mean_e ** 2
    var_m = (w * m ** 2).sum() / Z - mean_m ** 2
    return mean_e, mean_m, var_e, var_m
