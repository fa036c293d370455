"""Exact enumeration of the 2D Blume-Capel model on tiny periodic lattices.

Independent of the oracle: energies come from numpy rolls over row
configurations, not from the neighbour tables.  Used to pin the Metropolis
chain statistically (its stationary law must be exp(-beta E_J + lnW(D))).
"""
import itertools

import numpy as np


def row_configs(L):
    return np.array(list(itertools.product((-1, 0, 1), repeat=L)), dtype=np.int64)


def density_of_states(L, beta):
    """g[D, M + N] = sum over states with (D, M) of exp(-beta E_J), E_J = -sum_<ij> s_i s_j.

    Returns (g, logshift) with g scaled by exp(-logshift) to stay finite."""
    N = L * L
    R = row_configs(L)                     # (3^L, L)
    nr = len(R)
    intra = -(R * np.roll(R, -1, axis=1)).sum(1)          # horizontal bonds in a row
    inter = -(R[:, None, :] * R[None, :, :]).sum(2)       # vertical bonds between rows
    dr = (R * R).sum(1)
    mr = R.sum(1)
    emin = -2 * N
    shift = -beta * emin
    g = np.zeros((N + 1, 2 * N + 1), dtype=np.float64)
    assert L == 4, "enumeration written for four rows"
    i1, i2, i3 = np.meshgrid(np.arange(nr), np.arange(nr), np.arange(nr), indexing="ij")
    i1, i2, i3 = i1.ravel(), i2.ravel(), i3.ravel()
    base = intra[i1] + intra[i2] + intra[i3] + inter[i1, i2] + inter[i2, i3]
    dbase = dr[i1] + dr[i2] + dr[i3]
    mbase = mr[i1] + mr[i2] + mr[i3]
    for r0 in range(nr):
        E = base + intra[r0] + inter[r0, i1] + inter[i3, r0]
        D = dbase + dr[r0]
        M = mbase + mr[r0]
        w = np.exp(-beta * E - shift)
        g += np.bincount(D * (2 * N + 1) + M + N, weights=w,
                         minlength=g.size).reshape(g.shape)
    return g, shift
