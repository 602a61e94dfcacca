"""Overlapping multislab (oms) S-system, assembled on the MI355X.

Mirror of multislab/oms.py (SimonDirckx/SslabLU): slab (:17-52), omsStats (:53-59) and
oms (:61-320) with compute_global_dofs, compute_stmaps, construct_Stot_helper,
construct_Stot_and_rhstot_linearOperator, construct_Stot_and_rhstot and
construct_rhstot, and oms_lu (:321-525, the S maps left as factor-backed
operators).  Conventions kept: the S-system u_i + S_{i,i-1} u_{i-1} +
S_{i,i+1} u_{i+1} = rhs_i with S_J = (A_ii^{-1} A_ib[:, J])[Ic] and
rhs = -(A_ii^{-1} A_ib[:, Igb] f)[Ic]; S_rk_list = [[S_r], [S_l, S_r], ..., [S_l]];
contiguous global dof ranges per interface.

The reference discretizes every double slab on its own (oms.py:160-205); here the
double slabs of a dSlabs(N) decomposition of the unit square or cube with the 'hps' /
'hpsalt' discretization are assembled together on the GPU (every single slab once, see
sslablu_amd.hps and sslablu_amd.hps3d).  S blocks come back as DeviceMatrix (dense assembler) or HBSMAT
(rkHBS assembler) objects that live in HBM; rhs_list entries are numpy arrays.
"""
from __future__ import annotations

import time

import numpy as np
import torch
from scipy.sparse.linalg import LinearOperator

from ..hps import HPSSlabOperator, SlabGrid2D
from ..hps3d import HPSSlabOperator3D, SlabGrid3D
from ..linop import DeviceMatrix
from ..solver.solver import hps_grid_params, stMap


class slab:
    """One double slab: geometry and the global-boundary predicate (oms.py:17-52)."""

    def __init__(self, geom, gb_vec, transform=None):
        self.geom = geom
        self.transform = transform
        self.gb_vec = gb_vec
        try:
            ndim = np.asarray(self.geom).shape[-1]
            probe = np.linspace(0.1, 0.9, 5 * ndim).reshape(5, ndim)
            res = np.asarray(gb_vec(probe))
            assert res.ndim == 1 and res.shape[0] == 5
        except Exception:
            raise ValueError("Input gb needs to accept inputs of size numpoints x ndim")

    def compute_idxs_and_pts(self, solver):
        """Il, Ir (slab-edge points not on the global boundary), Ic (centre line), Igb
        (global boundary) and the interior / boundary coordinates (oms.py:38-52)."""
        XX = solver.XX
        XXb = XX[solver.Ib, ...]
        XXi = XX[solver.Ii, ...]
        xl = self.geom[0][0]
        xr = self.geom[1][0]
        xc = (xl + xr) / 2.0
        gbb = np.asarray(self.gb_vec(XXb), dtype=bool)
        Il = np.where((np.abs(XXb[..., 0] - xl) < 1e-14) & ~gbb)[0]
        Ir = np.where((np.abs(XXb[..., 0] - xr) < 1e-14) & ~gbb)[0]
        Ic = np.where(np.abs(XXi[..., 0] - xc) < 1e-14)[0]
        Igb = np.where(gbb)[0]
        return Il, Ir, Ic, Igb, XXi, XXb


class omsStats:
    def __init__(self):
        self.compression = None
        self.compr_timing = None
        self.discr_timing = None
        self.sampl_timing = None


class oms:
    """Overlapping multislab S-system (oms.py:61-320)."""

    def __init__(self, slabList, pdo, gb, solver_opts, connectivity):
        self.slabList = slabList
        self.pdo = pdo
        self.connectivity = connectivity
        self.opts = solver_opts
        self.gb = gb
        self.glob_target_dofs = []
        self.glob_source_dofs = []
        self.localSolver = None
        self.nbytes = 0
        self.densebytes = 0
        self.stats = omsStats()
        self.op = None

    # -- bookkeeping (oms.py:83-97) -----------------------------------------------------
    def compute_global_dofs(self):
        if not self.glob_source_dofs and self.glob_target_dofs:
            src = []
            for k in range(len(self.connectivity)):
                l, r = self.connectivity[k]
                if l < 0:
                    src.append([self.glob_target_dofs[r]])
                elif r < 0:
                    src.append([self.glob_target_dofs[l]])
                else:
                    src.append([self.glob_target_dofs[l], self.glob_target_dofs[r]])
            self.glob_source_dofs = src

    # -- discretization ------------------------------------------------------------------
    def _grid(self):
        slabs = [np.asarray(s, dtype=np.float64) for s in self.slabList]
        ndim = slabs[0].shape[1]
        periodic = self._check_chain(slabs)
        H = (slabs[0][1][0] - slabs[0][0][0]) / 2.0
        # the chain of dSlabs(N) has N - 1 double slabs from x0; the periodic one has N,
        # the first centred on x0 (geometry/geom_3D/squareTorus.py:268-277)
        N = len(slabs) if periodic else len(slabs) + 1
        x0 = slabs[0][0][0] + (H if periodic else 0.0)
        lo, hi = slabs[0][0], slabs[0][1]
        if ndim == 3:
            q, hx, hy, hz = hps_grid_params(self.opts, ndim=3)
            nl = (int(round(H / hx)), int(round((hi[1] - lo[1]) / hy)),
                  int(round((hi[2] - lo[2]) / hz)))
            return SlabGrid3D(N, *nl, q, origin=(x0, lo[1], lo[2]),
                              size=(N * H, hi[1] - lo[1], hi[2] - lo[2]), periodic=periodic)
        if ndim != 2:
            raise NotImplementedError("the GPU HPS discretization is 2-D or 3-D")
        q, hx, hy = hps_grid_params(self.opts)
        nx = int(round(H / hx))
        ny = int(round((hi[1] - lo[1]) / hy))
        return SlabGrid2D(N, nx, ny, q, x0=x0, width=N * H, y0=lo[1], height=hi[1] - lo[1],
                          periodic=periodic)

    def _spectral(self):
        """The 'spectral' discretization (solver/spectral/spectralSolver.py:198-235): one
        Chebyshev tensor grid of ord + 1 points per axis on every double slab; 2-D with
        variable coefficients (SpectralSlabs2D), 3-D with constant ones (SpectralSlabs3D,
        fast diagonalization)."""
        slabs = [np.asarray(s, dtype=np.float64) for s in self.slabList]
        if self._check_chain(slabs):
            raise NotImplementedError("periodic slabs use the 'hps' discretization")
        N = len(slabs) + 1
        H = (slabs[0][1][0] - slabs[0][0][0]) / 2.0
        lo, hi = slabs[0][0], slabs[0][1]
        ordv = [int(o) for o in np.atleast_1d(self.opts.ord)]
        if slabs[0].shape[1] == 2:
            from ..spectral2d import SpectralSlabs2D
            return SpectralSlabs2D(self.pdo, N, ordv[0], ordv[-1], x0=lo[0], width=N * H,
                                   y0=lo[1], height=hi[1] - lo[1]).build()
        from ..spectral3d import SpectralSlabs3D
        if np.any(np.abs(lo - np.array([0.0, 0.0, 0.0])) > 1e-12) or abs(N * H - 1) > 1e-12 \
                or np.any(np.abs(hi[1:] - 1.0) > 1e-12):
            raise NotImplementedError("3-D spectral slabs are assembled on the unit cube")
        P = np.random.default_rng(0).random((16, 3))
        coef = []
        for name, default in (("c11", 1.0), ("c22", 1.0), ("c33", 1.0), ("c", 0.0)):
            f = getattr(self.pdo, name, None)
            v = np.full(16, default) if f is None else np.asarray(f(P), dtype=np.float64)
            if np.ptp(v) > 1e-14 * max(1.0, np.abs(v).max()):
                raise NotImplementedError("3-D spectral slabs need constant coefficients")
            coef.append(float(v[0]))
        for name in ("c1", "c2", "c3", "c12", "c13", "c23"):
            if getattr(self.pdo, name, None) is not None:
                raise NotImplementedError("3-D spectral slabs: no first-order or mixed terms")
        op = SpectralSlabs3D(N, ordv[0] + 1, ordv[1] + 1, ordv[-1] + 1, *coef).build()
        op.SiM, op.SiP = op.blocks()
        return op

    def _check_chain(self, slabs):
        """The double slabs must be consecutive ([x + kH - H, x + kH + H]-style, equal
        cross sections) with the chain connectivity of dSlabs, or the periodic one of
        squareTorus.dSlabs; returns True for the periodic case."""
        n = len(slabs)
        first = slabs[0][0][0]
        H = (slabs[0][1][0] - slabs[0][0][0]) / 2.0
        for k, s in enumerate(slabs):
            if abs(s[0][0] - (first + k * H)) > 1e-12 or abs(s[1][0] - (first + (k + 2) * H)) > 1e-12 \
                    or np.any(np.abs(s[:, 1:] - slabs[0][:, 1:]) > 1e-12):
                raise NotImplementedError("slabs must be the consecutive double slabs of a "
                                          "box (geometry.geom_2D.square / geom_3D.cube / "
                                          "geom_3D.squareTorus dSlabs)")
        conn = [list(c) for c in self.connectivity]
        if n >= 3 and conn == [[(k - 1) % n, (k + 1) % n] for k in range(n)]:
            return True
        if conn != [[k - 1 if k > 0 else -1, k + 1 if k < n - 1 else -1] for k in range(n)]:
            raise NotImplementedError("connectivity must be the chain of dSlabs or the "
                                      "periodic ring of squareTorus.dSlabs")
        return False

    def discretize(self):
        """Assemble (once) the S maps of every double slab on the GPU (2-D: dense maps;
        3-D HPS: the pair factors behind the S-map operators)."""
        if self.op is None:
            t0 = time.time()
            if self.opts.type == "spectral":
                self.op = self._spectral()
            else:
                g = self._grid()
                cls = HPSSlabOperator3D if isinstance(g, SlabGrid3D) else HPSSlabOperator
                self.op = cls(self.pdo, g).build()
            torch.cuda.synchronize()
            self.stats.discr_timing = (time.time() - t0) / max(len(self.slabList), 1)
        return self.op

    def compute_stmaps(self, k):
        """stMap objects of the left / right S maps of double slab k (oms.py:99-129):
        dense device blocks in 2-D, device operators (sparse-solve-like applies through
        the pair factors) in 3-D, as the reference's LinearOperators."""
        op = self.discretize()
        if not hasattr(op, "g"):  # spectral slabs: centre line / plane of double slab k
            m = op.m
            XXc = op.interface_points()[k * m:(k + 1) * m]
            XXl, XXr = XXc.copy(), XXc.copy()
            XXl[:, 0] -= op.H
            XXr[:, 0] += op.H
            return stMap(op.SiM[k], XXl, XXc), stMap(op.SiP[k], XXr, XXc)
        g = op.g
        XXc = g.interface_points()[k * g.m:(k + 1) * g.m]
        XXl, XXr = XXc.copy(), XXc.copy()
        XXl[:, 0] -= g.H
        XXr[:, 0] += g.H
        if isinstance(g, SlabGrid3D):
            return stMap(op.smap(k, "l"), XXl, XXc), stMap(op.smap(k, "r"), XXr, XXc)
        return stMap(op.SiM[k], XXl, XXc), stMap(op.SiP[k], XXr, XXc)

    # -- S-system (oms.py:131-264) ------------------------------------------------------
    def construct_Stot_helper(self, bc, assembler, dbg=0):
        op = self.discretize()
        nc = op.g.m if hasattr(op, "g") else op.m
        S_rk_list, glob_target_dofs = [], []
        compress_t = sample_t = 0.0
        for k in range(len(self.slabList)):
            st_l, st_r = self.compute_stmaps(k)
            left, right = self.connectivity[k]
            blocks = []
            for st, present in ((st_l, left >= 0), (st_r, right >= 0)):
                if not present:
                    continue
                M = assembler.assemble(st, dbg=dbg)
                compress_t += assembler.stats.timeCompress
                sample_t += assembler.stats.timeSample
                self.nbytes += assembler.stats.nbytes
                self.densebytes += nc * nc * 8
                blocks.append(DeviceMatrix(M) if isinstance(M, torch.Tensor) else M)
            S_rk_list.append(blocks)
            glob_target_dofs.append(range(k * nc, (k + 1) * nc))
        rhs_list = []
        if bc is not None:
            rhs = self.rhs_device(bc)
            rhs_list = [rhs[k * nc:(k + 1) * nc, 0].cpu().numpy() for k in range(len(self.slabList))]
        nsl = max(len(self.connectivity) - 1, 1)
        self.stats.compression = self.nbytes / self.densebytes if self.densebytes else None
        self.stats.sampl_timing = sample_t / nsl
        self.stats.compr_timing = compress_t / nsl
        self.glob_target_dofs = glob_target_dofs
        self.glob_source_dofs = []
        self.compute_global_dofs()
        self.nc = nc
        return S_rk_list, rhs_list, nc * len(self.slabList), nc

    def boundary_points(self):
        """Global boundary nodes in the order expected by rhs_device (bottom, top,
        left, right)."""
        return self.discretize().g.gb_points()

    def rhs_device(self, bc):
        """S-system right-hand sides (N_tot x r device tensor) for boundary data bc:
        a callable of points (one column) or an array of values at boundary_points()."""
        op = self.discretize()
        if not hasattr(op, "g"):  # spectral slabs: every double slab has its own grid
            if not callable(bc):
                raise ValueError("spectral slabs take boundary data as a callable of points")
            return op.rhs_from_function(bc)
        F = bc(self.boundary_points()) if callable(bc) else bc
        if not isinstance(F, torch.Tensor):
            F = torch.as_tensor(np.asarray(F, dtype=np.float64))
        return op.rhs_from_boundary(F)

    def construct_Stot_and_rhstot_linearOperator(self, S_rk_list, rhs_list, Ntot, nc, dbg=0):
        """S = I + off-diagonal blocks as a LinearOperator, and the stacked rhs
        (oms.py:267-304)."""
        rhstot = np.zeros(shape=(Ntot,))
        for k, r in enumerate(rhs_list):
            rhstot[k * nc:(k + 1) * nc] = r

        def smatmat(v, transpose=False):
            vt = np.asarray(v, dtype=np.float64)
            v2 = vt[:, None] if vt.ndim == 1 else vt
            res = v2.copy()
            for i in range(len(self.glob_target_dofs)):
                for j in range(len(self.glob_source_dofs[i])):
                    B = S_rk_list[i][j]
                    if not transpose:
                        res[self.glob_target_dofs[i]] += np.asarray(B @ v2[self.glob_source_dofs[i][j]])
                    else:
                        res[self.glob_source_dofs[i][j]] += np.asarray(B.T @ v2[self.glob_target_dofs[i]])
            return res[:, 0] if vt.ndim == 1 else res

        Linop = LinearOperator(shape=(Ntot, Ntot), matvec=smatmat,
                               rmatvec=lambda v: smatmat(v, transpose=True), matmat=smatmat,
                               rmatmat=lambda v: smatmat(v, transpose=True), dtype=np.float64)
        return Linop, rhstot

    def construct_Stot_and_rhstot(self, bc, assembler, dbg=0):
        S_rk_list, rhs_list, Ntot, nc = self.construct_Stot_helper(bc, assembler, dbg)
        return self.construct_Stot_and_rhstot_linearOperator(S_rk_list, rhs_list, Ntot, nc, dbg)

    def construct_rhstot(self, bc):
        """Stacked S-system right-hand side for boundary data bc (numpy); the reference
        leaves this as a TODO (oms.py:316-320)."""
        return self.rhs_device(bc)[:, 0].cpu().numpy()


class oms_lu(oms):
    """Overlapping multislab S-system without interface compression (oms.py:321-525):
    the S maps stay operators backed by the local factorizations -- in 3-D the pair LU
    factors of HPSSlabOperator3D (PairSMap), in 2-D the dense maps the HPS pair solve
    produces (DeviceMatrix).  construct_Stot_helper(bc, dbg) takes no assembler."""

    def construct_Stot_helper(self, bc, dbg=0):
        op = self.discretize()
        nc = op.g.m if hasattr(op, "g") else op.m
        S_lu_list, glob_target_dofs = [], []
        for k in range(len(self.slabList)):
            st_l, st_r = self.compute_stmaps(k)
            left, right = self.connectivity[k]
            blocks = [st.A for st, present in ((st_l, left >= 0), (st_r, right >= 0)) if present]
            S_lu_list.append([DeviceMatrix(b) if isinstance(b, torch.Tensor) else b
                              for b in blocks])
            glob_target_dofs.append(range(k * nc, (k + 1) * nc))
            if dbg > 0:
                print("overlapping slab ", k + 1, " of ", len(self.connectivity), " done")
        rhs_list = []
        if bc is not None:
            rhs = self.rhs_device(bc)
            rhs_list = [rhs[k * nc:(k + 1) * nc, 0].cpu().numpy() for k in range(len(self.slabList))]
        self.glob_target_dofs = glob_target_dofs
        self.glob_source_dofs = []
        self.compute_global_dofs()
        self.nc = nc
        return S_lu_list, rhs_list, nc * len(self.slabList), nc

    def construct_Stot_and_rhstot(self, bc, assembler=None, dbg=0):
        S_lu_list, rhs_list, Ntot, nc = self.construct_Stot_helper(bc, dbg)
        return self.construct_Stot_and_rhstot_linearOperator(S_lu_list, rhs_list, Ntot, nc, dbg)
