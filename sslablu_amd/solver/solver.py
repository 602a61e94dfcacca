"""Local slab solver wrapper and the source-target map container.

Mirror of solver/solver.py (SimonDirckx/SslabLU): stMap (:30-36), solverOptions
(:39-50) and solverWrapper (:65-163).  The discretization types 'hps' and 'hpsalt' are
this project's Chebyshev HPS on the GPU, 2-D (sslablu_amd.hps) and 3-D
(sslablu_amd.hps3d) (the reference's HPS submodules are not part of it): ord = [q]*d
Chebyshev points per leaf side (q = p + 2, the reference drivers' p_disc) and a the
leaf half-sizes, as in the reference drivers (test/3D/cubeTest.py:114-118).

solverWrapper.construct(geom, PDE) discretizes one double slab and exposes the
reference's attributes on its skeleton: XX (node coordinates), Ib (the slab boundary,
ordered left, right, bottom, top), Ii (the interior skeleton: the centre line), plus
smap(J) = (A_ii^{-1} A_ib[:, J])[Ic] as device tensors.  The skeleton system is the
flux balance of the two single slabs' DtN maps T^a, T^b on the centre line c:
    A_ii = T^a_cc + T^b_cc,   A_ib = [T^a_cb | T^b_cb]   (columns in Ib order),
the static condensation of the HPS leaves onto it (the reference's hpsalt condenses onto
all leaf faces; the maps it yields, smap and the rhs, are the same).  solver_ii is
A_ii^{-1} as a linear operator (its LU, batched-LU kernels; modes N and T), Aib the
coupling block; in 2-D also Aii, Abi and Abb (the boundary rows of the same condensed
system).  ('spectral' slabs expose smap only.)
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
    """type: 'hps' (GPU Chebyshev HPS; 'hpsalt', the name the reference's 3-D drivers use,
    selects the same discretization, with ord = p_disc points per leaf side; a: leaf
    half-sizes) or 'spectral' (one Chebyshev grid per double slab, ord = the polynomial
    degree per axis: ord + 1 points)."""

    def __init__(self, type: str, ord, a=None):
        self.type = type
        self.ord = ord
        self.a = a


def hps_grid_params(opts, *_, ndim=2):
    """(q, hx, hy) of an 'hps' / 'hpsalt' solverOptions, or (q, hx, hy, hz) for ndim 3."""
    if opts.type not in ("hps", "hpsalt"):
        raise NotImplementedError(f"discretization {opts.type!r} is not provided on the GPU "
                                  "(available: 'hps', 'hpsalt')")
    ordv = np.atleast_1d(np.asarray(opts.ord, dtype=np.int64))
    q = int(ordv[0])
    if np.any(ordv != q):
        raise NotImplementedError("equal leaf orders along every axis are required")
    a = np.asarray(opts.a, dtype=np.float64).ravel()
    if a.size == 1:
        a = np.full(ndim, a[0])
    if a.size < ndim:
        raise ValueError(f"solverOptions.a needs {ndim} leaf half-sizes")
    return (q, *(2.0 * a[:ndim]))


class solverWrapper:
    def __init__(self, opts: solverOptions):
        self.ord = opts.ord
        self.type = opts.type
        self.a = opts.a
        self.opts = opts
        self.constructed = False

    def construct(self, geom, PDE, verbose=False, compute_inverse=True):
        geom = np.asarray(geom, dtype=np.float64)
        self.ndim = geom.shape[1]
        if self.type == "spectral":
            return self._construct_spectral(geom, PDE)
        if self.ndim == 3:
            return self._construct3(geom, PDE)
        if self.ndim != 2:
            raise NotImplementedError("the GPU HPS discretization is 2-D or 3-D")
        from ..hps import HPSSlabOperator, SlabGrid2D
        (xl, yl), (xr, yr) = geom
        q, hx, hy = hps_grid_params(self.opts)
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
        return self._finish(centre.shape[0])

    def _construct3(self, geom, PDE):
        """One 3-D double slab: two single slabs of HPS leaves (sslablu_amd.hps3d); the
        boundary is ordered [left face | right face | side faces of the first single slab
        | side faces of the second], the interior skeleton is the centre plane."""
        from ..hps3d import HPSSlabOperator3D, SlabGrid3D
        lo, hi = geom
        q, hx, hy, hz = hps_grid_params(self.opts, ndim=3)
        H = (hi[0] - lo[0]) / 2.0
        nl = [int(round(H / hx)), int(round((hi[1] - lo[1]) / hy)),
              int(round((hi[2] - lo[2]) / hz))]
        grid = SlabGrid3D(2, *nl, q, origin=lo, size=hi - lo)
        self.op = HPSSlabOperator3D(PDE, grid).build()
        g = grid
        pts = g.gb_points()  # [sides of slab 0 | sides of slab 1 | left | right]
        ns, m = g.n_side, g.m
        centre = g.interface_points()
        self.XX = np.concatenate([pts[2 * ns:2 * ns + m], pts[2 * ns + m:], pts[:2 * ns], centre])
        return self._finish(centre.shape[0])

    def _construct_spectral(self, geom, PDE):
        """One 2-D double slab on a spectral grid of (ord[0]+1) x (ord[1]+1) Chebyshev
        points (solver/spectral/spectralSolver.py:198-235); the boundary is ordered
        [left | right | bottom | top] (corners dropped), the skeleton is the centre line."""
        if self.ndim != 2:
            raise NotImplementedError("solverWrapper('spectral') on one double slab is 2-D "
                                      "(3-D spectral slabs are assembled through oms)")
        from ..spectral2d import SpectralSlabs2D
        (xl, yl), (xr, yr) = geom
        ordv = [int(o) for o in np.atleast_1d(self.opts.ord)]
        self.op = SpectralSlabs2D(PDE, 2, ordv[0], ordv[-1], x0=xl, width=xr - xl, y0=yl,
                                  height=yr - yl).build()
        self.XX = np.concatenate([self.op.boundary_points(0), self.op.interface_points()])
        return self._finish(self.op.m)

    def _finish(self, ncentre):
        nb = self.XX.shape[0] - ncentre
        self.Ib = np.arange(nb)
        self.Ii = np.arange(nb, self.XX.shape[0])
        self.XXb = self.XX[self.Ib]
        self.XXi = self.XX[self.Ii]
        self.ndofs = self.XX.shape[0]
        self.constructed = True
        return self

    def smap(self, which):
        """'l', 'r' or 'gb': the S map from the left face, the right face, or the rest of
        the boundary (the columns of A_ii^{-1} A_ib restricted to the centre skeleton).
        2-D: dense device tensors; 3-D: device operators (matmat / rmatmat)."""
        if which not in ("l", "r", "gb"):
            raise ValueError(which)
        if self.ndim == 3:
            from ..hps3d import PairSMap, SideMap
            return SideMap(self.op) if which == "gb" else PairSMap(self.op, 0, which)
        Y = self.op.Y[0]
        m = self.op.m if self.type == "spectral" else self.op.g.m
        if which == "l":
            return Y[:, :m]
        if which == "r":
            return Y[:, m:2 * m]
        return Y[:, 2 * m:]

    # -- the condensed skeleton system (solver/solver.py:100-144 attributes) ------------
    def _skeleton2d(self):
        """Aii, Aib, Abi, Abb of a 2-D HPS double slab from its single slabs' DtN maps."""
        if getattr(self, "_sk", None) is None:
            from ..hps import pair_maps
            g = self.op.g
            dev = self.op.device
            Ta, Tb = self.op.Ts[0], self.op.Ts[1]
            ca, cb, cola, colb = (torch.tensor(x, device=dev)  # copies the cached maps
                                  for x in pair_maps(g.nx, g.ny, g.p))
            ina, inb = cola >= 0, colb >= 0
            ja, jb = cola.clamp(min=0), colb.clamp(min=0)
            Aii = Ta[ca][:, ca] + Tb[cb][:, cb]
            Aib = Ta[ca][:, ja] * ina + Tb[cb][:, jb] * inb
            Abi = Ta[ja][:, ca] * ina[:, None] + Tb[jb][:, cb] * inb[:, None]
            Abb = (Ta[ja][:, ja] * (ina[:, None] & ina[None, :])
                   + Tb[jb][:, jb] * (inb[:, None] & inb[None, :]))
            self._sk = (Aii.contiguous(), Aib.contiguous(), Abi.contiguous(), Abb.contiguous())
        return self._sk

    def _require_hps(self):
        if not self.constructed:
            raise ValueError("solverWrapper: call construct first")
        if self.type == "spectral":
            raise NotImplementedError("'spectral' slabs expose the S maps (smap) only")

    @property
    def Aii(self):
        self._require_hps()
        if self.ndim == 3:
            raise NotImplementedError("3-D: A_ii is kept only as its LU (solver_ii)")
        return self._skeleton2d()[0]

    @property
    def Aib(self):
        self._require_hps()
        if self.ndim == 3:
            op = self.op
            return torch.cat([op.Bl[0], op.Br[0], op.G[0]], dim=1)
        return self._skeleton2d()[1]

    @property
    def Abi(self):
        self._require_hps()
        if self.ndim == 3:
            raise NotImplementedError("3-D: only A_ii^{-1} and A_ib are kept")
        return self._skeleton2d()[2]

    @property
    def Abb(self):
        self._require_hps()
        if self.ndim == 3:
            raise NotImplementedError("3-D: only A_ii^{-1} and A_ib are kept")
        return self._skeleton2d()[3]

    @property
    def solver_ii(self):
        """A_ii^{-1} of the condensed skeleton system as a linear operator
        (solver/solver.py:117-118, 137-140; solver_Aii of the reference's HPS solvers)."""
        self._require_hps()
        if getattr(self, "_sii", None) is None:
            if self.ndim == 3:
                op = self.op
                self._sii = InteriorSolve(None, None, inv=op.tinv[0])
            else:
                self._sii = InteriorSolve.factor(self._skeleton2d()[0])
        return self._sii


def _solve_dir(self, uu_dir, ff_body=None):
    """u_i = A_ii^{-1} (ff_body - A_ib u_b) on the skeleton (solver/pde_solver.py:169-201):
    uu_dir (n_b,) or (n_b, r) in Ib order, ff_body (n_i,) / (n_i, r) or None; numpy in ->
    numpy out."""
    was_np = not isinstance(uu_dir, torch.Tensor)
    Aib = self.Aib
    u = torch.as_tensor(np.asarray(uu_dir, dtype=np.float64)) if was_np else uu_dir
    u = u.to(Aib.device, torch.float64)
    one = u.dim() == 1
    U = u[:, None] if one else u
    rhs = -(Aib @ U)
    if ff_body is not None:
        f = torch.as_tensor(np.asarray(ff_body, dtype=np.float64)) if not isinstance(
            ff_body, torch.Tensor) else ff_body
        f = f.to(Aib.device, torch.float64)
        rhs = rhs + (f[:, None] if f.dim() == 1 else f)
    X = self.solver_ii @ rhs
    X = X[:, 0] if one else X
    return X.cpu().numpy() if was_np else X


solverWrapper.solve_dir = _solve_dir


class InteriorSolve:
    """X -> A^{-1} X (mode N) or A^{-T} X (mode T) from batched-LU factors of one m x m
    matrix (sslab_dgetrf_batched / sslab_dgetrs_batched / sslab_dgetrs_trans_batched), or
    from its explicit inverse (inv: one GEMM per apply); numpy in -> numpy out, device
    tensors in -> device tensors out, (m,) or (m, r)."""

    device_op = True

    def __init__(self, lu, piv, mode="N", inv=None):
        self.lu, self.piv, self.inv = lu, piv, inv
        m = (inv if inv is not None else lu).shape[0]
        self.shape = (m, m)
        self.dtype = np.float64
        self.mode = mode

    @staticmethod
    def factor(A):
        from .._lib import call, ptr, stream_handle
        m = A.shape[0]
        lu = A.detach().to(torch.float64).clone().contiguous()
        piv = torch.empty(m, dtype=torch.int32, device=lu.device)
        info = torch.zeros(1, dtype=torch.int32, device=lu.device)
        call("sslab_dgetrf_batched", m, ptr(lu), m, m * m, ptr(piv), m, ptr(info), 1,
             stream_handle())
        if int(info.item()) != 0:
            import warnings
            warnings.warn("solver_ii: singular A_ii", RuntimeWarning)
        return InteriorSolve(lu, piv)

    @property
    def T(self):
        return InteriorSolve(self.lu, self.piv, "T" if self.mode == "N" else "N", self.inv)

    def _apply(self, v, trans):
        from .._lib import call, ptr, stream_handle
        was_np = not isinstance(v, torch.Tensor)
        x = torch.as_tensor(np.asarray(v, dtype=np.float64)) if was_np else v
        x = x.to((self.inv if self.inv is not None else self.lu).device, torch.float64)
        one = x.dim() == 1
        B = (x[:, None] if one else x).contiguous()
        m, r = B.shape
        if m != self.shape[0]:
            raise ValueError(f"solver_ii: {self.shape[0]} rows expected, got {m}")
        if self.inv is not None:
            from ..hbs import bgemm
            X = bgemm(self.inv[None], B[None], ta=trans)[0] if r else torch.empty_like(B)
            X = X[:, 0] if one else X
            return X.cpu().numpy() if was_np else X
        X = torch.empty_like(B)
        perm = torch.empty(m, dtype=torch.int32, device=B.device)
        s = stream_handle()
        if r:
            if trans:
                W = torch.empty_like(B)
                call("sslab_dgetrs_trans_batched", m, r, ptr(self.lu), m, m * m, ptr(self.piv),
                     m, ptr(B), r, m * r, ptr(X), r, m * r, ptr(W), ptr(perm), 1, s)
            else:
                call("sslab_dgetrs_batched", m, r, ptr(self.lu), m, m * m, ptr(self.piv), m,
                     ptr(B), r, m * r, ptr(X), r, m * r, ptr(perm), 1, s)
        X = X[:, 0] if one else X
        return X.cpu().numpy() if was_np else X

    def matmat(self, v):
        return self._apply(v, self.mode == "T")

    def rmatmat(self, v):
        return self._apply(v, self.mode == "N")

    matvec = matmat
    rmatvec = rmatmat

    def __matmul__(self, v):
        return self.matmat(v)
