"""Oracle: 3-D Chebyshev tensor-grid (spectral) slab discretization and its oms S-system
(numpy FP64, CPU, small grids only).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates
  * the Chebyshev points / differentiation matrices and the 3-D operator
    L = -C11 Dxx(x)I(x)I - C22 I(x)Dyy(x)I - C33 I(x)I(x)Dzz + C  on the x-major, then y, then z
    node order -- solver/spectral/spectralSolver.py:132-195 (cheb, spectralD, spectralD2,
    constructPDO3D) and :226-251 (the 3-D grid, Ji = interior, Jx = boundary nodes);
  * the double slabs of the unit cube -- geometry/geom_3D/cube.py:32-46 (dSlabs);
  * the S maps S_J = (A_ii^{-1} A_ib[:, J])[Ic] and rhs = -(A_ii^{-1} A_ib[:, Igb] f)[Ic]
    -- multislab/oms.py:38-52, 99-129, 180-186,
by dense assembly and dense solves (no diagonalization), so it checks the device
module's fast-diagonalization formulas independently.

Parity pins (tests/test_oracle_spectral3d.py): the Chebyshev second derivative is exact
on polynomials; the S-system solved densely reproduces a harmonic polynomial exactly
(u = x^2 - y^2 + 3 z x is reproduced to round-off: it lies in the discrete space) and
u = exp(sqrt(2) pi x) sin(pi y) sin(pi z) to spectral accuracy.
"""
from __future__ import annotations

import numpy as np


def cheb_points_d2(q: int, a: float, b: float):
    """q Chebyshev points ascending on [a, b] and the second-derivative matrix
    (spectralSolver.py:132-160)."""
    p = q - 1
    x = np.cos(np.pi * np.arange(q) / p)
    c = np.ones(q)
    c[0] = c[-1] = 2.0
    c *= (-1.0) ** np.arange(q)
    D = np.outer(c, 1.0 / c) / (x[:, None] - x[None, :] + np.eye(q))
    D -= np.diag(D.sum(axis=1))
    pts = (x - x[0]) * (b - a) / (x[-1] - x[0]) + a
    D = D / ((pts[-1] - pts[0]) / (x[-1] - x[0]))
    return pts, D @ D


def slab_operator(x, y, z, c11=1.0, c22=1.0, c33=1.0, c=0.0):
    """Dense 3-D operator on the tensor grid (x-major node order) and the node list
    (constructPDO3D, spectralSolver.py:176-195, constant coefficients)."""
    (xp, Dxx), (yp, Dyy), (zp, Dzz) = x, y, z
    Ex, Ey, Ez = np.eye(len(xp)), np.eye(len(yp)), np.eye(len(zp))
    L = -c11 * np.kron(np.kron(Dxx, Ey), Ez)
    L -= c22 * np.kron(np.kron(Ex, Dyy), Ez)
    L -= c33 * np.kron(np.kron(Ex, Ey), Dzz)
    L += c * np.eye(L.shape[0])
    XX = np.stack([np.kron(np.kron(xp, np.ones_like(yp)), np.ones_like(zp)),
                   np.kron(np.kron(np.ones_like(xp), yp), np.ones_like(zp)),
                   np.kron(np.kron(np.ones_like(xp), np.ones_like(yp)), zp)], axis=1)
    return L, XX


def oms_spectral3d(N, nx, ny, nz, g, c11=1.0, c22=1.0, c33=1.0, c=0.0):
    """S-system of dSlabs(N) on the unit cube (cube.py:32-46), every double slab one
    nx x ny x nz Chebyshev grid.  Returns (SiM, SiP, rhs, interface_points) with blocks
    (nb, m, m) (SiM[0] = SiP[-1] = 0), rhs (nb*m, r) for Dirichlet data g(P) -> (n, r)."""
    H = 1.0 / N
    nb = N - 1
    rhs, pts, SiM, SiP = [], [], [], []
    tol = 1e-12
    for i in range(nb):
        xc = (i + 1) * H
        L, XX = slab_operator(cheb_points_d2(nx, xc - H, xc + H), cheb_points_d2(ny, 0.0, 1.0),
                              cheb_points_d2(nz, 0.0, 1.0), c11, c22, c33, c)
        x, yy, zz = XX[:, 0], XX[:, 1], XX[:, 2]
        inner_yz = (yy > tol) & (yy < 1 - tol) & (zz > tol) & (zz < 1 - tol)
        Ji = np.where((x > xc - H + tol) & (x < xc + H - tol) & inner_yz)[0]
        Jx = np.setdiff1d(np.arange(len(x)), Ji)
        Ic = np.where((np.abs(x - xc) < tol) & inner_yz)[0]
        Il = np.where((np.abs(x - (xc - H)) < tol) & inner_yz)[0]
        Ir = np.where((np.abs(x - (xc + H)) < tol) & inner_yz)[0]
        gb = Jx[(x[Jx] < tol) | (x[Jx] > 1 - tol) | ~inner_yz[Jx]]  # nodes on the cube boundary
        Aii = L[np.ix_(Ji, Ji)]
        pos = {j: k for k, j in enumerate(Ji)}
        ic = [pos[j] for j in Ic]
        Sl = np.linalg.solve(Aii, L[np.ix_(Ji, Il)])[ic]
        Sr = np.linalg.solve(Aii, L[np.ix_(Ji, Ir)])[ic]
        f = np.asarray(g(XX[gb]), dtype=np.float64).reshape(len(gb), -1)
        rhs.append(-np.linalg.solve(Aii, L[np.ix_(Ji, gb)] @ f)[ic])
        SiM.append(Sl if i > 0 else np.zeros_like(Sl))
        SiP.append(Sr if i < nb - 1 else np.zeros_like(Sr))
        pts.append(XX[Ic])
    return np.stack(SiM), np.stack(SiP), np.concatenate(rhs), np.concatenate(pts)
