"""Oracle: the 2-D non-overlapping multislab (SOMS) system of a box tiling (numpy FP64,
CPU).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates test/rkprop/SOMS.py (SimonDirckx/SslabLU):
  * SOMS_solver(px, py, nbx, nby, Lx, Ly, kh) (:9-169): the rectangle [0, Lx] x [0, Ly]
    tiled by nbx x nby boxes, each carrying px+1 x py+1 Chebyshev points; the unknowns are
    the interior Chebyshev nodes of every box edge.  For every interior edge the two boxes
    sharing it are joined into one box with a finer Chebyshev grid across the joint
    ((3p)//2 rounded down to even, so the shared edge is a grid line); the constant-
    coefficient operator -u_xx - u_yy - kh^2 u is discretized on it, its boundary values
    are glued from the pair's outer edge unknowns (the two boxes' edge nodes along the
    joined sides interpolated to the joined grid, gluing_mat :197-223), and the edge's row
    of Stot is  u_edge + (L_ii^{-1} L_ib[:, outer] C)[centre] u_outer = 0;
  * the unknown ordering (:98-113): per box column, its vertical edges bottom to top,
    then its horizontal edges bottom to top; finally the vertical edges of x = Lx;
  * the joined-box layout of joined (:171-195) and the solve of the tests,
    u_i = -S_ii^{-1} S_ib u_b (test/rkprop/SOMStest2.py:58-69).

Parity pins (tests/test_oracle_soms2d.py): a harmonic cubic is reproduced exactly
(Laplace); the gluing map interpolates polynomials exactly; Helmholtz plane waves converge
spectrally; the unknowns' coordinates are distinct and lie on the box edges; a planted
sign error in one family of rows breaks the exactness.
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.chebyshev as chebpoly


def cheb_grid(n, length):
    """n+1 ascending Chebyshev points on [0, length] and the derivative matrix on them
    (Trefethen's cheb, mapped by t = (1 - x) length / 2)."""
    x = np.cos(np.pi * np.arange(n + 1) / n)
    c = np.ones(n + 1)
    c[0] = c[-1] = 2.0
    c *= (-1.0) ** np.arange(n + 1)
    D = np.outer(c, 1.0 / c) / (x[:, None] - x[None, :] + np.eye(n + 1))
    D -= np.diag(D.sum(axis=1))
    return (1.0 - x) * length / 2.0, -2.0 * D / length


def joined_layout(a, b):
    """Tensor grid a x b (index ia * len(b) + ib): interior nodes, boundary nodes, and,
    as positions in the boundary list, its four sides without corners (first coordinate
    low / high, second low / high) and, as positions in the interior list, the centre line
    of the first coordinate (SOMS.py:171-195)."""
    A = np.repeat(a, len(b))
    B = np.tile(b, len(a))
    la, lb = a[-1] - a[0], b[-1] - b[0]
    tol = 1e-10
    inner = (A > tol) & (A < la - tol) & (B > tol) & (B < lb - tol)
    Ii = np.flatnonzero(inner)
    Ib = np.flatnonzero(~inner)
    Ab, Bb = A[Ib], B[Ib]
    mid_a = (Ab > tol) & (Ab < la - tol)
    mid_b = (Bb > tol) & (Bb < lb - tol)
    sides = [np.flatnonzero((np.abs(Ab) < tol) & mid_b), np.flatnonzero((np.abs(Ab - la) < tol) & mid_b),
             np.flatnonzero((np.abs(Bb) < tol) & mid_a), np.flatnonzero((np.abs(Bb - lb) < tol) & mid_a)]
    Ic = np.flatnonzero(np.abs(A[Ii] - la / 2) < tol)
    return Ii, Ib, sides, Ic


def glue(t, t_joined, n_other):
    """Gluing map of a joined pair (SOMS.py:197-223): columns [first side (n_other - 2) |
    low side of box 1, box 2 | high side of box 1, box 2 | last side], rows the joined
    box's [first side | low side | high side | last side] (corners dropped).  Along the
    joint the two boxes' edge nodes are fitted by a least-squares Chebyshev series of
    degree (joined interior points - 1) (SVD pseudo-inverse, singular values below 1e-15
    relative dropped) and evaluated at the joined interior points."""
    L = t[-1] - t[0]
    src = np.concatenate([t[1:-1], L + t[1:-1]])
    dst = t_joined[1:-1]
    deg = len(dst) - 1
    Vs = chebpoly.chebvander((src - L) / L, deg)
    Vd = chebpoly.chebvander((dst - L) / L, deg)
    U, sv, Vt = np.linalg.svd(Vs, full_matrices=False)
    k = int(np.sum(sv > 1e-15 * sv[0]))
    cc = Vd @ (Vt[:k].T / sv[:k]) @ U[:, :k].T          # (len(dst), 2 (len(t) - 2))
    e, h, w = n_other - 2, len(dst), len(t) - 2
    C = np.zeros((2 * e + 2 * h, 2 * e + 4 * w))
    C[:e, :e] = np.eye(e)
    C[e:e + h, e:e + 2 * w] = cc
    C[e + h:e + 2 * h, e + 2 * w:e + 4 * w] = cc
    C[e + 2 * h:, e + 4 * w:] = np.eye(e)
    return C


def pair_map(t_along, t_other, L_along, kh):
    """S of one joined pair: rows the centre line of the joined box, columns the glued
    outer unknowns; S = (L_ii^{-1} L_ib[:, sides]) [centre] C (the negative of the S of
    SOMS.py:84-87, which the assembly negates back)."""
    n_al = len(t_along) - 1
    nj = (3 * n_al) // 2
    nj -= nj % 2
    tj, Dj = cheb_grid(nj, 2 * L_along)
    _, Do = cheb_grid(len(t_other) - 1, t_other[-1] - t_other[0])
    no = len(t_other)
    Lop = (-np.kron(Dj @ Dj, np.eye(no)) - np.kron(np.eye(len(tj)), Do @ Do)
           - kh * kh * np.eye(len(tj) * no))
    Ii, Ib, (s_lo, s_hi, o_lo, o_hi), Ic = joined_layout(tj, t_other)
    outer = np.concatenate([s_lo, o_lo, o_hi, s_hi])
    X = np.linalg.solve(Lop[np.ix_(Ii, Ii)], Lop[np.ix_(Ii, Ib[outer])])
    return X[Ic] @ glue(t_along, tj, no)


class Layout:
    """Unknown numbering of SOMS.py:98-113 for nbx x nby boxes with ex = px - 1 interior
    nodes per horizontal edge and ey = py - 1 per vertical edge."""

    def __init__(self, nbx, nby, ex, ey):
        self.nbx, self.nby, self.ex, self.ey = nbx, nby, ex, ey
        self.col = nby * ey + (nby + 1) * ex           # unknowns per box column

    def vert(self, ix, iy):
        """Vertical edge x = ix hx, y in [iy hy, (iy+1) hy]."""
        return ix * self.col + iy * self.ey

    def horz(self, ix, iy):
        """Horizontal edge y = iy hy, x in [ix hx, (ix+1) hx]."""
        return ix * self.col + self.nby * self.ey + iy * self.ex

    @property
    def n(self):
        return self.nbx * self.col + self.nby * self.ey


def soms_system(px, py, nbx, nby, Lx=1.0, Ly=1.0, kh=0.0, _plant=None):
    """(Stot, XYtot, Ii, Ib) of SOMS_solver (SOMS.py:9-169)."""
    hx, hy = Lx / nbx, Ly / nby
    tx, _ = cheb_grid(px, hx)
    ty, _ = cheb_grid(py, hy)
    ex, ey = len(tx) - 2, len(ty) - 2
    lay = Layout(nbx, nby, ex, ey)
    XY = np.zeros((lay.n, 2))
    for ix in range(nbx + 1):
        for iy in range(nby):
            i = lay.vert(ix, iy)
            XY[i:i + ey] = np.stack([np.full(ey, ix * hx), iy * hy + ty[1:-1]], 1)
    for ix in range(nbx):
        for iy in range(nby + 1):
            i = lay.horz(ix, iy)
            XY[i:i + ex] = np.stack([ix * hx + tx[1:-1], np.full(ex, iy * hy)], 1)
    Sx = pair_map(tx, ty, hx, kh)     # joined across a vertical edge (boxes left / right)
    Sy = pair_map(ty, tx, hy, kh)     # joined across a horizontal edge (boxes below / above)
    if _plant == "flip_x":
        Sx = -Sx
    S = np.eye(lay.n)
    rng_ = lambda a, k: np.arange(a, a + k)  # noqa: E731
    for ix in range(nbx):              # horizontal interior edges
        for iy in range(1, nby):
            cols = np.concatenate([rng_(lay.horz(ix, iy - 1), ex),
                                   rng_(lay.vert(ix, iy - 1), 2 * ey),
                                   rng_(lay.vert(ix + 1, iy - 1), 2 * ey),
                                   rng_(lay.horz(ix, iy + 1), ex)])
            S[np.ix_(rng_(lay.horz(ix, iy), ex), cols)] = Sy
    for ix in range(1, nbx):           # vertical interior edges
        for iy in range(nby):
            cols = np.concatenate([rng_(lay.vert(ix - 1, iy), ey),
                                   rng_(lay.horz(ix - 1, iy), ex), rng_(lay.horz(ix, iy), ex),
                                   rng_(lay.horz(ix - 1, iy + 1), ex), rng_(lay.horz(ix, iy + 1), ex),
                                   rng_(lay.vert(ix + 1, iy), ey)])
            S[np.ix_(rng_(lay.vert(ix, iy), ey), cols)] = Sx
    tol = 1e-14
    x, y = XY[:, 0], XY[:, 1]
    on_b = (np.abs(x) < tol) | (np.abs(x - Lx) < tol) | (np.abs(y) < tol) | (np.abs(y - Ly) < tol)
    Ib = np.flatnonzero(on_b)
    Ii = np.flatnonzero(~on_b)
    return S, XY, Ii, Ib


def soms_solve(S, Ii, Ib, ub):
    """u_i = -S_ii^{-1} S_ib u_b (test/rkprop/SOMStest2.py:60-68)."""
    return -np.linalg.solve(S[np.ix_(Ii, Ii)], S[np.ix_(Ii, Ib)] @ ub)
