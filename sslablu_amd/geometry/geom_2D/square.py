"""Unit square and its double-slab decomposition.

Mirror of geometry/geom_2D/square.py (SimonDirckx/SslabLU): bnds, gb_np / gb (the global
boundary predicate, :6-19), box_geom, bounds and dSlabs(N) (:33-47): N-1 double slabs
[(n+1)H - H, (n+1)H + H] x [0, 1] with connectivity [[-1,1],[0,2],...,[N-3,-1]].
"""
import numpy as np

bnds = [[0, 0], [1, 1]]


def gb_np(p):
    p = np.asarray(p)
    return ((np.abs(p[:, 0] - bnds[0][0]) < 1e-14) | (np.abs(p[:, 0] - bnds[1][0]) < 1e-14)
            | (np.abs(p[:, 1] - bnds[0][1]) < 1e-14) | (np.abs(p[:, 1] - bnds[1][1]) < 1e-14))


def gb(p, jax_avail=False, torch_avail=False):
    return gb_np(p)


def box_geom(jax_avail=False, torch_avail=True):
    return np.array(bnds)


def bounds():
    return bnds


def dSlabs(N):
    H = (bnds[1][0] - bnds[0][0]) / N
    slabs, connectivity = [], []
    for n in range(N - 1):
        c = bnds[0][0] + (n + 1) * H
        slabs.append([[c - H, bnds[0][1]], [c + H, bnds[1][1]]])
        if n == 0:
            connectivity.append([-1, 1])
        elif n == N - 2:
            connectivity.append([n - 1, -1])
        else:
            connectivity.append([n - 1, n + 1])
    return slabs, connectivity, H
