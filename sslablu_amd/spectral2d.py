"""Spectral double slabs in 2-D with variable coefficients, on the GPU.

The reference's 'spectral' slab discretization (solver/spectral/spectralSolver.py:132-173,
198-235; solver/solver.py:145-157): every double slab [x_c - H, x_c + H] x [y0, y0 + Hy]
is one tensor Chebyshev grid of (ordx + 1) x (ordy + 1) points (ordx even, so the centre
line x = x_c is a grid column), with the collocation operator
    A[u] = -c11 u_xx - c22 u_yy + c1 u_x + c2 u_y + c u         (pdo/pdo.py:19-39)
at the interior nodes.  The oms maps (multislab/oms.py:99-129, 180-186) are the rows of
the centre line of A_ii^{-1} A_ib: S_l from the left edge, S_r from the right edge, and
the bottom / top (and, at the ends, the domain's left / right) boundary-data maps.  One
batched call (sslab_spectral2d_slab_maps) assembles every slab's A_ii / A_ib, factors
A_ii (batched LU) and solves for all boundary columns.

Interface nodes are the ordy - 1 interior Chebyshev points in y, the same on every
interface, so neighbouring slabs share their unknowns.  Difference from the reference, on
purpose: constructPDO2D (spectralSolver.py:167-170) multiplies c1 into the y-derivative
and drops c2; here c1 multiplies u_x and c2 u_y, as pdo/pdo.py defines the operator.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, sizes2, stream_handle
from .hps import cheb_unit

F64 = torch.float64
I32 = torch.int32


def _coef(fn, P):
    if fn is None:
        return None
    v = fn(P)
    if isinstance(v, torch.Tensor):
        v = v.detach().cpu().numpy()
    v = np.asarray(v, dtype=np.float64)
    if v.ndim == 0:
        v = np.full(P.shape[0], float(v))
    return v.reshape(-1)


class SpectralSlabs2D:
    """S-system of a 2-D PDO on the rectangle [x0, x0 + W] x [y0, y0 + Hy] cut into N
    single slabs (N - 1 double slabs / interfaces), spectral double slabs of
    (ordx + 1) x (ordy + 1) Chebyshev points."""

    def __init__(self, pdo, N, ordx, ordy, x0=0.0, width=1.0, y0=0.0, height=1.0, device=None):
        if N < 2:
            raise ValueError("need at least two single slabs (one interface)")
        if ordx < 2 or ordx % 2:
            raise ValueError("ordx must be even (the centre line is a grid column)")
        if ordy < 2:
            raise ValueError("ordy >= 2")
        _lib.require_gpu()
        _lib.load()
        self.pdo = pdo
        self.N, self.nb = N, N - 1
        self.qx, self.qy = ordx + 1, ordy + 1
        self.m = self.qy - 2
        self.x0, self.width, self.y0, self.height = x0, width, y0, height
        self.H = width / N
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        tx, Dx = cheb_unit(self.qx)
        ty, Dy = cheb_unit(self.qy)
        self.tx, self.ty = tx, ty
        hx, hy = 2 * self.H, height
        self._D = [Dx / hx, (Dx @ Dx) / hx ** 2, Dy / hy, (Dy @ Dy) / hy ** 2]
        self.Y = None

    def _dev(self, a, dtype=F64):
        return torch.as_tensor(np.ascontiguousarray(a)).to(device=self.device, dtype=dtype)

    def slab_x(self, k):
        """The qx grid abscissae of double slab k."""
        return self.x0 + k * self.H + self.tx * 2 * self.H

    def y_points(self):
        return self.y0 + self.ty * self.height

    def interior_points(self, k):
        """(qx-2)(qy-2) interior nodes of double slab k, row (a-1)(qy-2) + (b-1)."""
        X, Yg = np.meshgrid(self.slab_x(k)[1:-1], self.y_points()[1:-1], indexing="ij")
        return np.stack([X.ravel(), Yg.ravel()], 1)

    def boundary_points(self, k):
        """Boundary nodes of double slab k in column order [left | right | bottom | top]
        (corners dropped)."""
        xs, ys = self.slab_x(k), self.y_points()
        xi, yi = xs[1:-1], ys[1:-1]
        left = np.stack([np.full_like(yi, xs[0]), yi], 1)
        right = np.stack([np.full_like(yi, xs[-1]), yi], 1)
        bottom = np.stack([xi, np.full_like(xi, ys[0])], 1)
        top = np.stack([xi, np.full_like(xi, ys[-1])], 1)
        return np.concatenate([left, right, bottom, top])

    def interface_points(self):
        yi = self.y_points()[1:-1]
        return np.concatenate([np.stack([np.full_like(yi, self.x0 + (k + 1) * self.H), yi], 1)
                               for k in range(self.nb)])

    def build(self, stream=None):
        nb, qx, qy = self.nb, self.qx, self.qy
        P = np.concatenate([self.interior_points(k) for k in range(nb)])
        coefs = [_coef(getattr(self.pdo, name, None), P) for name in ("c11", "c22", "c1", "c2", "c")]
        cdev = [self._dev(c) if c is not None else None for c in coefs]
        D = [self._dev(d) for d in self._D]
        nbd = 2 * (qy - 2) + 2 * (qx - 2)
        self.Y = torch.empty(nb, self.m, nbd, dtype=F64, device=self.device)
        wd, wi = sizes2("sslab_spectral2d_sizes", nb, qx, qy)
        work = torch.empty(wd, dtype=F64, device=self.device)
        iwork = torch.empty(wi, dtype=I32, device=self.device)
        info = torch.zeros(nb, dtype=I32, device=self.device)
        call("sslab_spectral2d_slab_maps", nb, qx, qy, *[ptr(d) for d in D],
             *[ptr(c) if c is not None else None for c in cdev], ptr(self.Y), ptr(work),
             ptr(iwork), ptr(info), stream_handle(stream))
        self.info_bad = int((info != 0).sum().item())
        m = self.m
        self.SiM = self.Y[:, :, :m].contiguous()
        self.SiP = self.Y[:, :, m:2 * m].contiguous()
        if self.info_bad:
            import warnings
            warnings.warn(f"{self.info_bad} singular slab solve(s) in the spectral assembly",
                          RuntimeWarning)
        return self

    def S_rk_list(self):
        nb = self.nb
        return [[self.SiP[0]]] + [[self.SiM[i], self.SiP[i]] for i in range(1, nb - 1)] + \
               [[self.SiM[nb - 1]]]

    def rhs_from_function(self, bc) -> torch.Tensor:
        """S-system right-hand sides rhs_k = -(A_ii^{-1} A_ib[:, Igb] f)[Ic] (oms.py:
        180-186) for Dirichlet data bc: points (n, 2) -> (n,) or (n, r)."""
        from .hbs import bgemm
        nb, m = self.nb, self.m
        cols = []
        for k in range(nb):
            v = np.asarray(bc(self.boundary_points(k)), dtype=np.float64)
            v = v[:, None] if v.ndim == 1 else v
            if k > 0:
                v[:m] = 0.0              # the left edge is interface k - 1, not boundary
            if k < nb - 1:
                v[m:2 * m] = 0.0         # the right edge is interface k + 1
            cols.append(v)
        F = self._dev(np.stack(cols))
        return (-bgemm(self.Y, F)).reshape(nb * m, -1)
