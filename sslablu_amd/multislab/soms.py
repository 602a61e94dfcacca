"""The 2-D non-overlapping multislab (SOMS) system of a box tiling, on the MI355X.

Mirror of test/rkprop/SOMS.py (SimonDirckx/SslabLU): SOMS_solver(px, py, nbx, nby, Lx=1,
Ly=1, kh=0.) (:9-169) returning (Stot, XYtot, Ii, Ib), and the solve of its tests
u_i = -S_ii^{-1} S_ib u_b (test/rkprop/SOMStest2.py:58-69).

The rectangle [0, Lx] x [0, Ly] is tiled by nbx x nby boxes of (px + 1) x (py + 1)
Chebyshev points; the unknowns are the interior nodes of every box edge, numbered per box
column (its vertical edges bottom to top, then its horizontal edges bottom to top), then
the vertical edges of x = Lx (:98-113).  Every interior edge's rows read the outer edges
of the two boxes sharing it, through the map of the joined pair: the pair discretized as
one box with a finer Chebyshev grid across the joint ((3p)//2, even), the constant-
coefficient operator -u_xx - u_yy - kh^2 u on it, its boundary values glued from the outer
edges' nodes (least-squares Chebyshev fit along the joint, gluing_mat :197-223):
    row block = (L_ii^{-1} L_ib[:, outer] C)[centre line].
Two such maps exist (pairs joined across a vertical or a horizontal edge).

On the device: the joined boxes' interior operators are factored by the batched LU, the
maps formed by the LU solve and one GEMM with the gluing matrix; Stot is kept block
sparse (one gather and one batched GEMM with the shared map per edge family apply it) and
formed densely on request; the solve factors S_ii by the batched LU (dense, for up to
solve_dense_max unknowns) or runs the device GMRES on the block-sparse apply.  The small
constant operators (Chebyshev matrices, the Kronecker-sum operator of the joined box,
the gluing matrix) are prepared on the host.
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.chebyshev as chebpoly
import torch

from .. import _lib
from .._lib import call, ptr, stream_handle

F64 = torch.float64
I32 = torch.int32


def _cheb(n, length):
    x = np.cos(np.pi * np.arange(n + 1) / n)
    c = np.ones(n + 1)
    c[0] = c[-1] = 2.0
    c *= (-1.0) ** np.arange(n + 1)
    D = np.outer(c, 1.0 / c) / (x[:, None] - x[None, :] + np.eye(n + 1))
    D -= np.diag(D.sum(axis=1))
    return (1.0 - x) * length / 2.0, -2.0 * D / length


def _glue(t, tj, n_other):
    L = t[-1] - t[0]
    src = np.concatenate([t[1:-1], L + t[1:-1]])
    dst = tj[1:-1]
    deg = len(dst) - 1
    U, sv, Vt = np.linalg.svd(chebpoly.chebvander((src - L) / L, deg), full_matrices=False)
    k = int(np.sum(sv > 1e-15 * sv[0]))
    cc = chebpoly.chebvander((dst - L) / L, deg) @ (Vt[:k].T / sv[:k]) @ U[:, :k].T
    e, h, w = n_other - 2, len(dst), len(t) - 2
    C = np.zeros((2 * e + 2 * h, 2 * e + 4 * w))
    C[:e, :e] = np.eye(e)
    C[e:e + h, e:e + 2 * w] = cc
    C[e + h:e + 2 * h, e + 2 * w:e + 4 * w] = cc
    C[e + 2 * h:, e + 4 * w:] = np.eye(e)
    return C


def _lu(A):
    n = A.shape[-1]
    nb = A.shape[0]
    piv = torch.empty(nb, n, dtype=I32, device=A.device)
    info = torch.zeros(nb, dtype=I32, device=A.device)
    call("sslab_dgetrf_batched", n, ptr(A), n, n * n, ptr(piv), n, ptr(info), nb, stream_handle())
    return piv, info


def _lu_solve(lu, piv, B):
    nb, n, r = B.shape
    X = torch.empty_like(B)
    perm = torch.empty(nb, n, dtype=I32, device=B.device)
    call("sslab_dgetrs_batched", n, r, ptr(lu), n, n * n, ptr(piv), n, ptr(B), r, n * r, ptr(X),
         r, n * r, ptr(perm), nb, stream_handle())
    return X


class SOMS2D:
    """SOMS system of an nbx x nby box tiling; build() forms the two pair maps on the GPU."""

    solve_dense_max = 20000

    def __init__(self, px, py, nbx, nby, Lx=1.0, Ly=1.0, kh=0.0, device=None):
        _lib.require_gpu()
        _lib.load()
        if px < 2 or py < 2 or nbx < 1 or nby < 1:
            raise ValueError("SOMS2D: px, py >= 2 and at least one box per direction")
        self.px, self.py, self.nbx, self.nby = int(px), int(py), int(nbx), int(nby)
        self.Lx, self.Ly, self.kh = float(Lx), float(Ly), float(kh)
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.hx, self.hy = self.Lx / nbx, self.Ly / nby
        self.tx, _ = _cheb(self.px, self.hx)
        self.ty, _ = _cheb(self.py, self.hy)
        self.ex, self.ey = len(self.tx) - 2, len(self.ty) - 2
        self.col = nby * self.ey + (nby + 1) * self.ex
        self.n = nbx * self.col + nby * self.ey
        self._fac = None

    # -- numbering (SOMS.py:98-113) -------------------------------------------------------
    def vert(self, ix, iy):
        return ix * self.col + iy * self.ey

    def horz(self, ix, iy):
        return ix * self.col + self.nby * self.ey + iy * self.ex

    def points(self):
        XY = np.zeros((self.n, 2))
        for ix in range(self.nbx + 1):
            for iy in range(self.nby):
                i = self.vert(ix, iy)
                XY[i:i + self.ey, 0] = ix * self.hx
                XY[i:i + self.ey, 1] = iy * self.hy + self.ty[1:-1]
        for ix in range(self.nbx):
            for iy in range(self.nby + 1):
                i = self.horz(ix, iy)
                XY[i:i + self.ex, 0] = ix * self.hx + self.tx[1:-1]
                XY[i:i + self.ex, 1] = iy * self.hy
        return XY

    def index_sets(self):
        XY = self.points()
        tol = 1e-14
        x, y = XY[:, 0], XY[:, 1]
        on_b = ((np.abs(x) < tol) | (np.abs(x - self.Lx) < tol) | (np.abs(y) < tol)
                | (np.abs(y - self.Ly) < tol))
        return XY, np.flatnonzero(~on_b), np.flatnonzero(on_b)

    def _edge_blocks(self):
        """Row starts and column index lists of the two edge families."""
        r = lambda a, k: np.arange(a, a + k)  # noqa: E731
        ex, ey = self.ex, self.ey
        hr, hc, vr, vc = [], [], [], []
        for ix in range(self.nbx):
            for iy in range(1, self.nby):
                hr.append(r(self.horz(ix, iy), ex))
                hc.append(np.concatenate([r(self.horz(ix, iy - 1), ex), r(self.vert(ix, iy - 1), 2 * ey),
                                          r(self.vert(ix + 1, iy - 1), 2 * ey),
                                          r(self.horz(ix, iy + 1), ex)]))
        for ix in range(1, self.nbx):
            for iy in range(self.nby):
                vr.append(r(self.vert(ix, iy), ey))
                vc.append(np.concatenate([r(self.vert(ix - 1, iy), ey),
                                          r(self.horz(ix - 1, iy), ex), r(self.horz(ix, iy), ex),
                                          r(self.horz(ix - 1, iy + 1), ex),
                                          r(self.horz(ix, iy + 1), ex),
                                          r(self.vert(ix + 1, iy), ey)]))
        return hr, hc, vr, vc

    # -- the pair maps (SOMS.py:27-95) ------------------------------------------------------
    def _pair_operator(self, t_along, t_other, L_along):
        n_al = len(t_along) - 1
        nj = (3 * n_al) // 2
        nj -= nj % 2
        tj, Dj = _cheb(nj, 2 * L_along)
        _, Do = _cheb(len(t_other) - 1, t_other[-1] - t_other[0])
        no = len(t_other)
        Lop = (-np.kron(Dj @ Dj, np.eye(no)) - np.kron(np.eye(len(tj)), Do @ Do)
               - self.kh ** 2 * np.eye(len(tj) * no))
        A = np.repeat(tj, no)
        B = np.tile(t_other, len(tj))
        la, lb, tol = tj[-1], t_other[-1] - t_other[0], 1e-10
        inner = (A > tol) & (A < la - tol) & (B > tol) & (B < lb - tol)
        Ii, Ib = np.flatnonzero(inner), np.flatnonzero(~inner)
        Ab, Bb = A[Ib], B[Ib]
        mid_a = (Ab > tol) & (Ab < la - tol)
        mid_b = (Bb > tol) & (Bb < lb - tol)
        outer = np.concatenate([np.flatnonzero((np.abs(Ab) < tol) & mid_b),
                                np.flatnonzero((np.abs(Bb) < tol) & mid_a),
                                np.flatnonzero((np.abs(Bb - lb) < tol) & mid_a),
                                np.flatnonzero((np.abs(Ab - la) < tol) & mid_b)])
        Ic = np.flatnonzero(np.abs(A[Ii] - la / 2) < tol)
        return (Lop[np.ix_(Ii, Ii)], Lop[np.ix_(Ii, Ib[outer])], Ic,
                _glue(t_along, tj, no))

    def build(self):
        dev = self.device
        maps = []
        for t_al, t_ot, L_al in ((self.tx, self.ty, self.hx), (self.ty, self.tx, self.hy)):
            Lii, Lib, Ic, C = self._pair_operator(t_al, t_ot, L_al)
            lu = torch.tensor(Lii[None], device=dev)
            piv, info = _lu(lu)
            X = _lu_solve(lu, piv, torch.tensor(Lib[None], device=dev))[0]
            Xc = X.index_select(0, torch.as_tensor(Ic, device=dev)).contiguous()
            from ..ops import dgemm
            maps.append(dgemm(Xc, torch.tensor(C, device=dev)))
            if int(info.item()) != 0:
                import warnings
                warnings.warn("SOMS2D: singular joined-box operator (kh at a resonance?)",
                              RuntimeWarning)
        self.Sx, self.Sy = maps
        hr, hc, vr, vc = self._edge_blocks()
        t = lambda a: torch.as_tensor(np.array(a, dtype=np.int64), device=dev)  # noqa: E731
        self._h = (t(hr), t(hc)) if hr else None
        self._v = (t(vr), t(vc)) if vr else None
        return self

    # -- Stot ---------------------------------------------------------------------------
    def apply(self, X):
        """Stot X for an (n, r) device panel (block-sparse: identity plus the edge rows)."""
        Y = X.clone()
        for fam, S in ((self._h, self.Sy), (self._v, self.Sx)):
            if fam is None:
                continue
            rows, cols = fam
            nbk, k = cols.shape
            Xg = X.index_select(0, cols.reshape(-1)).view(nbk, k, -1)
            from ..hbs import bgemm
            Z = bgemm(S[None], Xg)                          # (nbk, e, r)
            Y.index_add_(0, rows.reshape(-1), Z.reshape(-1, X.shape[1]))
        return Y

    def dense(self):
        S = torch.eye(self.n, dtype=F64, device=self.device)
        for fam, B in ((self._h, self.Sy), (self._v, self.Sx)):
            if fam is None:
                continue
            rows, cols = fam
            S[rows[:, :, None], cols[:, None, :]] = B.expand(rows.shape[0], *B.shape)
        return S

    def solve(self, ub, method=None, tol=1e-12):
        """u_i = -S_ii^{-1} S_ib u_b for boundary values ub (len(Ib),) or (len(Ib), r) in
        Ib order: dense LU of S_ii (kept for later calls) up to solve_dense_max interior
        unknowns, else GMRES on the block-sparse apply (method 'dense' / 'gmres' forces
        one).  numpy in -> numpy out."""
        XY, Ii, Ib = self.index_sets()
        was_np = not isinstance(ub, torch.Tensor)
        u = torch.as_tensor(np.asarray(ub, dtype=np.float64)) if was_np else ub
        u = u.to(self.device, F64)
        one = u.dim() == 1
        U = u[:, None] if one else u
        dev = self.device
        Iit, Ibt = torch.as_tensor(Ii, device=dev), torch.as_tensor(Ib, device=dev)
        method = method or ("dense" if len(Ii) <= self.solve_dense_max else "gmres")
        Full = torch.zeros(self.n, U.shape[1], dtype=F64, device=dev)
        Full[Ibt] = U
        rhs = -self.apply(Full).index_select(0, Iit)     # -S_ib u_b (Full is 0 on Ii)
        if method == "dense":
            if self._fac is None:
                Sii = self.dense()[Iit][:, Iit].contiguous()[None]
                piv, info = _lu(Sii)
                self._fac = (Sii, piv)
            lu, piv = self._fac
            X = _lu_solve(lu, piv, rhs[None].contiguous())[0]
        else:
            from ..krylov import gmres

            def A(V):
                F = torch.zeros(self.n, V.shape[1], dtype=F64, device=dev)
                F[Iit] = V
                return self.apply(F).index_select(0, Iit)
            X, info = gmres(A, rhs.contiguous(), tol=tol, restart=100, maxiter=1000)
            self.last_gmres = info
        X = X[:, 0] if one else X
        return X.cpu().numpy() if was_np else X


def SOMS_solver(px, py, nbx, nby, Lx=1, Ly=1, kh=0.):
    """(Stot, XYtot, Ii, Ib) as test/rkprop/SOMS.py:9-169 returns them; Stot is a dense
    float64 device tensor (SOMS2D keeps the block-sparse form)."""
    op = SOMS2D(px, py, nbx, nby, Lx, Ly, kh).build()
    XY, Ii, Ib = op.index_sets()
    return op.dense(), XY, Ii, Ib
