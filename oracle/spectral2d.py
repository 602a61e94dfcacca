"""Oracle: 2-D spectral double slabs with variable coefficients (numpy/scipy FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates solver/spectral/spectralSolver.py:132-173, 198-235 (cheb, spectralD2,
constructPDO2D, spectralSolver) with multislab/oms.py:38-52, 99-129, 180-186 on top: per
double slab one (ordx+1) x (ordy+1) Chebyshev grid, L = -C11 (D2x (x) I) - C22 (I (x) D2y)
+ C1 (Dx (x) I) + C2 (I (x) Dy) + C (diagonal coefficient scalings), A_ii = L[Ji, Ji],
A_ib = L[Ji, Jx], S_J = (A_ii^{-1} A_ib[:, J])[Ic], rhs = -(A_ii^{-1} A_ib[:, Igb] f)[Ic].
The first-order terms follow pdo/pdo.py (c1 u_x + c2 u_y); constructPDO2D itself
multiplies c1 into the y-derivative and drops c2 (spectralSolver.py:167-170).

Pinned by tests/test_oracle_spectral2d.py: exactness on a harmonic polynomial of degree
<= ord, spectral convergence to exp(x) sin(y), the Laplace row-sum invariant, and a
planted sign error.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .hps import cheb_diff, eval_coef


def cheb_points(n, lo, hi):
    """n + 1 Chebyshev points ascending on [lo, hi] and d/dx on them
    (spectralSolver.py:132-151, 212-219)."""
    x, D = cheb_diff(n)
    t = (x - x[0]) / (x[-1] - x[0])          # 0 .. 1 ascending
    return lo + t * (hi - lo), D / ((hi - lo) / (x[-1] - x[0]))


def slab_operator(pdo, bounds, ordx, ordy, _plant=None):
    (xl, yl), (xr, yr) = bounds
    xp, Dx = cheb_points(ordx, xl, xr)
    yp, Dy = cheb_points(ordy, yl, yr)
    X, Yg = np.meshgrid(xp, yp, indexing="ij")
    XX = np.stack([X.ravel(), Yg.ravel()], 1)
    N = XX.shape[0]
    Ix, Iy = sp.identity(len(xp)), sp.identity(len(yp))
    diag = lambda f: sp.diags(eval_coef(f, XX)) if f is not None else None  # noqa: E731
    L = -diag(pdo.c11) @ sp.kron(Dx @ Dx, Iy) - diag(pdo.c22) @ sp.kron(Ix, Dy @ Dy)
    if getattr(pdo, "c1", None) is not None:
        L = L + diag(pdo.c1) @ sp.kron(Dx, Iy)
    if getattr(pdo, "c2", None) is not None:
        L = L + diag(pdo.c2) @ sp.kron(Ix, Dy)
    if getattr(pdo, "c", None) is not None:
        L = L + (-1.0 if _plant == "c_sign" else 1.0) * diag(pdo.c)
    L = L.tocsr()
    on_b = (np.abs(XX[:, 0] - xl) < 1e-14) | (np.abs(XX[:, 0] - xr) < 1e-14) | \
           (np.abs(XX[:, 1] - yl) < 1e-14) | (np.abs(XX[:, 1] - yr) < 1e-14)
    Ji, Jx = np.where(~on_b)[0], np.where(on_b)[0]
    return XX, Ji, Jx, L[Ji][:, Ji].tocsc(), L[Ji][:, Jx].tocsc()


def oms_stot(pdo, N, ordx, ordy, bc=None, _plant=None):
    """S_rk_list / rhs of the oms S-system on [0,1]^2 with N single slabs and spectral
    double slabs (geometry/geom_2D/square.py:33-47)."""
    H = 1.0 / N
    gb = lambda P: (np.abs(P[:, 0]) < 1e-14) | (np.abs(P[:, 0] - 1) < 1e-14) | \
        (np.abs(P[:, 1]) < 1e-14) | (np.abs(P[:, 1] - 1) < 1e-14)  # noqa: E731
    S_rk_list, rhs_list, Ic_pts, R_list = [], [], [], []
    for k in range(N - 1):
        xc = (k + 1) * H
        XX, Ji, Jx, Aii, Aib = slab_operator(pdo, ((xc - H, 0.0), (xc + H, 1.0)), ordx, ordy,
                                             _plant)
        XXb, XXi = XX[Jx], XX[Ji]
        gbb = gb(XXb)
        Il = np.where((np.abs(XXb[:, 0] - (xc - H)) < 1e-14) & ~gbb)[0]
        Ir = np.where((np.abs(XXb[:, 0] - (xc + H)) < 1e-14) & ~gbb)[0]
        Ic = np.where(np.abs(XXi[:, 0] - xc) < 1e-14)[0]
        Igb = np.where(gbb)[0]
        lu = spla.splu(Aii)
        smap = lambda J: lu.solve(Aib[:, J].toarray())[Ic]  # noqa: E731
        blocks = ([smap(Il)] if k > 0 else []) + ([smap(Ir)] if k < N - 2 else [])
        S_rk_list.append(blocks)
        R = smap(Igb)
        R_list.append(R)
        Ic_pts.append(XXi[Ic])
        if bc is not None:
            rhs_list.append(-R @ bc(XXb[Igb]))
    return dict(S_rk_list=S_rk_list, rhs_list=rhs_list, Ic_pts=Ic_pts, R_list=R_list,
                nc=len(Ic_pts[0]))
