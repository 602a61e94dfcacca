"""Unit cube and its double-slab decomposition.

Mirror of geometry/geom_3D/cube.py (SimonDirckx/SslabLU): bnds, gb_np / gb (the global
boundary predicate, :7-19; jax is not part of this stack, so the jax flag returns the
numpy predicate as well), box_geom, bounds and dSlabs(N) (:32-46): N-1 double slabs
[(n+1)H - H, (n+1)H + H] x [0, 1] x [0, 1] with connectivity [[-1,1],[0,2],...,[N-3,-1]].
"""
import numpy as np

bnds = [[0, 0, 0], [1, 1, 1]]


def gb_np(p):
    p = np.asarray(p)
    lo, hi = np.asarray(bnds[0], dtype=float), np.asarray(bnds[1], dtype=float)
    return np.any((np.abs(p - lo) < 1e-14) | (np.abs(p - hi) < 1e-14), axis=1)


def gb(p, jax_avail=False, torch_avail=False):
    if hasattr(p, "detach"):
        p = p.detach().cpu().numpy()
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
        slabs.append([[c - H, bnds[0][1], bnds[0][2]], [c + H, bnds[1][1], bnds[1][2]]])
        connectivity.append([-1 if n == 0 else n - 1, -1 if n == N - 2 else n + 1])
    return slabs, connectivity, H
