"""Local slab solver wrapper and the source-target map container.

Mirror of solver/solver.py (SimonDirckx/SslabLU): stMap (:30-36), solverOptions
(:39-50) and solverWrapper (:65-163).  The discretization type 'hps' is this project's
2-D Chebyshev HPS on the GPU (the reference's HPS submodules are not part of it):
ord = [q, q] Chebyshev points per leaf side (q = p + 2) and a = [hx/2, hy/2] the leaf
half-sizes, as in the reference drivers (test/3D/cubeTest.py:114-118).

solverWrapper.construct(geom, PDE) discretizes one double slab and exposes the
reference's attributes on its skeleton: XX (node coordinates), Ib (the slab boundary,
ordered left, right, bottom, top), Ii (the interior skeleton: the centre line), plus
smap(J) = (A_ii^{-1} A_ib[:, J])[Ic] as device tensors.  solver_ii (the full interior
solve) belongs to interior reconstruction, which is not provided.
"""
from __future__ import annotations

import numpy as np
import torch


class stMap:
    def __init__(self, A, XXI, XXJ, m_large=0, n_large=0):
        self.XXI = XXI
        self.XXJ = XXJ
        self.A = A
        self.m_large = m_large
        self.n_large = n_large


class solverOptions:
    """type: 'hps' (GPU Chebyshev HPS, 2-D); ord: points per leaf side; a: leaf half-sizes."""

    def __init__(self, type: str, ord, a=None):
        self.type = type
        self.ord = ord
        self.a = a


def hps_grid_params(opts, xl, xr, yl=0.0, yr=1.0):
    """(q, hx, hy) of an 'hps' solverOptions."""
    if opts.type != "hps":
        raise NotImplementedError(f"discretization {opts.type!r} is not provided on the GPU "
                                  "(available: 'hps')")
    q = int(opts.ord[0])
    if len(opts.ord) > 1 and int(opts.ord[1]) != q:
        raise NotImplementedError("equal leaf orders in x and y are required")
    a = np.asarray(opts.a, dtype=np.float64).ravel()
    if a.size == 1:
        a = np.array([a[0], a[0]])
    return q, 2 * a[0], 2 * a[1]


class solverWrapper:
    def __init__(self, opts: solverOptions):
        self.ord = opts.ord
        self.type = opts.type
        self.a = opts.a
        self.opts = opts
        self.constructed = False

    def construct(self, geom, PDE, verbose=False, compute_inverse=True):
        from ..hps import HPSSlabOperator, SlabGrid2D
        geom = np.asarray(geom, dtype=np.float64)
        self.ndim = geom.shape[1]
        if self.ndim != 2:
            raise NotImplementedError("the GPU HPS discretization is 2-D")
        (xl, yl), (xr, yr) = geom
        q, hx, hy = hps_grid_params(self.opts, xl, xr, yl, yr)
        H = (xr - xl) / 2.0
        nx = int(round(H / hx))
        ny = int(round((yr - yl) / hy))
        grid = SlabGrid2D(2, nx, ny, q, x0=xl, width=2 * H, y0=yl, height=yr - yl)
        self.op = HPSSlabOperator(PDE, grid).build()
        g = grid
        pts_b = g.gb_points()  # [bottom | top | left | right]
        nxp = 2 * g.nx * g.p
        bottom, top = pts_b[:nxp], pts_b[nxp:2 * nxp]
        left, right = pts_b[2 * nxp:2 * nxp + g.m], pts_b[2 * nxp + g.m:]
        centre = g.interface_points()
        self.XX = np.concatenate([left, right, bottom, top, centre])
        nb = self.XX.shape[0] - centre.shape[0]
        self.Ib = np.arange(nb)
        self.Ii = np.arange(nb, self.XX.shape[0])
        self.XXb = self.XX[self.Ib]
        self.XXi = self.XX[self.Ii]
        self.ndofs = self.XX.shape[0]
        self.constructed = True
        return self

    def smap(self, which):
        """'l', 'r' or 'gb': the S map from the left edge, the right edge, or the
        top/bottom edges (the columns of A_ii^{-1} A_ib restricted to the centre line)."""
        Y = self.op.Y[0]
        m = self.op.g.m
        if which == "l":
            return Y[:, :m]
        if which == "r":
            return Y[:, m:2 * m]
        if which == "gb":
            return Y[:, 2 * m:]
        raise ValueError(which)

    @property
    def solver_ii(self):
        raise NotImplementedError("interior solves (solution reconstruction inside slabs) are "
                                  "not provided; use the S maps (smap)")
