"""Multislab S-system built slab by slab through the local solver wrapper, on the MI355X.

Mirror of multislab/noms.py (SimonDirckx/SslabLU): slab (:17-43), omsStats (:44-48) and
oms (:50-260) with compute_global_dofs, compute_stmaps (:76-103: S_J = (A_ii^{-1}
A_ib[:, J])[Ic] as operators on solver_ii and Aib), construct_Stot_and_rhstot
(:105-253: per double slab solverWrapper.construct, the index sets, rhs =
-(A_ii^{-1} A_ib[:, Igb] f)[Ic], the assembler on both S maps, S_rk_list, the stacked rhs
and the S-system as a LinearOperator) and construct_rhstot (:255-260, a stub returning 0
in the reference, kept).

Where sslablu_amd.multislab.oms assembles all double slabs of a dSlabs decomposition
together, this module follows the reference's loop: any slab list and connectivity,
each double slab discretized by solverWrapper (2-D or 3-D HPS on the GPU) and its maps
applied through the wrapper's solver_ii (the LU of the condensed skeleton system) and
Aib, as device operators the assemblers sketch or apply to the identity.
"""
from __future__ import annotations

import time

import numpy as np
import torch
from scipy.sparse.linalg import LinearOperator

from ..linop import DeviceMatrix
from ..solver import solver as solverWrap
from ..solver.solver import stMap
from .oms import slab  # the same slab (geometry, global-boundary predicate, index sets)

F64 = torch.float64


class omsStats:
    def __init__(self):
        self.compression = None
        self.compr_timing = None
        self.discr_timing = None


class _SlabMap:
    """(A_ii^{-1} A_ib[:, J])[I] of one solverWrapper as a device operator."""

    device_op = True

    def __init__(self, solver, I, J):
        self.solver = solver
        dev = solver.Aib.device
        self.I = torch.as_tensor(np.asarray(I), device=dev)
        self.J = torch.as_tensor(np.asarray(J), device=dev)
        self.shape = (len(I), len(J))
        self.dtype = np.float64
        self._AibJ = solver.Aib.index_select(1, self.J).contiguous()

    def matmat(self, V):
        X = self.solver.solver_ii @ (self._AibJ @ V.to(F64))
        return X.index_select(0, self.I)

    def rmatmat(self, V):
        n = self.solver.solver_ii.shape[0]
        W = torch.zeros(n, V.shape[1], dtype=F64, device=V.device)
        W[self.I] = V.to(F64)
        return self._AibJ.t() @ (self.solver.solver_ii.T @ W)

    matvec = matmat
    rmatvec = rmatmat


class oms:
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

    def compute_global_dofs(self):
        if not self.glob_source_dofs and self.glob_target_dofs:
            out = []
            for k in range(len(self.connectivity)):
                l, r = self.connectivity[k][0], self.connectivity[k][1]
                if l < 0:
                    out.append([self.glob_target_dofs[r]])
                elif r < 0:
                    out.append([self.glob_target_dofs[l]])
                else:
                    out.append([self.glob_target_dofs[l], self.glob_target_dofs[r]])
            self.glob_source_dofs = out

    def compute_stmaps(self, Il, Ic, Ir, XXi, XXb, solver):
        st_r = stMap(_SlabMap(solver, Ic, Ir), XXb[Ir, ...], XXi[Ic, ...])
        st_l = stMap(_SlabMap(solver, Ic, Il), XXb[Il, ...], XXi[Ic, ...])
        return st_l, st_r

    def construct_Stot_and_rhstot(self, bc, assembler, dbg=0):
        Ntot = 0
        S_rk_list, rhs_list, glob_target_dofs = [], [], []
        start_c = 0
        discr, compr = 0.0, 0.0
        for k in range(len(self.slabList)):
            geom = np.array(self.slabList[k])
            sl = slab(geom, self.gb)
            t0 = time.time()
            solver = solverWrap.solverWrapper(self.opts)
            solver.construct(geom, self.pdo, verbose=dbg)
            discr += time.time() - t0
            Il, Ir, Ic, Igb, XXi, XXb = sl.compute_idxs_and_pts(solver)
            nc = len(Ic)
            self.nc = nc
            Ntot += nc
            glob_target_dofs.append(range(start_c, start_c + nc))
            start_c += nc
            fgb = torch.as_tensor(np.asarray(bc(XXb[Igb, ...]), dtype=np.float64),
                                  device=solver.Aib.device)
            rhs = -(solver.solver_ii @ (solver.Aib[:, torch.as_tensor(Igb, device=fgb.device)]
                                        @ fgb)).index_select(0, torch.as_tensor(Ic,
                                                                                device=fgb.device))
            rhs_list.append(rhs.cpu().numpy())
            st_l, st_r = self.compute_stmaps(Il, Ic, Ir, XXi, XXb, solver)
            t0 = time.time()
            mats = []
            for st, keep in ((st_l, len(Il) > 0), (st_r, len(Ir) > 0)):
                if keep:
                    M = assembler.assemble(st, dbg=dbg)
                    mats.append(DeviceMatrix(M) if isinstance(M, torch.Tensor) else M)
                    self.nbytes += assembler.stats.nbytes
                self.densebytes += int(np.prod(st.A.shape)) * 8
            compr += time.time() - t0
            if self.connectivity[k][0] < 0:
                S_rk_list.append([mats[-1]])
            elif self.connectivity[k][1] < 0:
                S_rk_list.append([mats[0]])
            else:
                S_rk_list.append(mats)
        n = max(len(self.connectivity) - 1, 1)
        self.stats.compression = self.nbytes / max(self.densebytes, 1)
        self.stats.compr_timing = compr / n
        self.stats.discr_timing = discr / n
        self.glob_target_dofs = glob_target_dofs
        self.compute_global_dofs()
        rhstot = np.concatenate(rhs_list)

        def smatmat(v, transpose=False):
            vt = np.asarray(v, dtype=np.float64)
            v2 = vt[:, None] if vt.ndim == 1 else vt
            res = v2.copy()
            for i in range(len(self.glob_target_dofs)):
                for j in range(len(self.glob_source_dofs[i])):
                    B = S_rk_list[i][j]
                    if not transpose:
                        res[self.glob_target_dofs[i]] += np.asarray(
                            B @ v2[self.glob_source_dofs[i][j]])
                    else:
                        res[self.glob_source_dofs[i][j]] += np.asarray(
                            B.T @ v2[self.glob_target_dofs[i]])
            return res[:, 0] if vt.ndim == 1 else res

        Linop = LinearOperator(shape=(Ntot, Ntot), matvec=smatmat,
                               rmatvec=lambda v: smatmat(v, transpose=True), matmat=smatmat,
                               rmatmat=lambda v: smatmat(v, transpose=True), dtype=np.float64)
        self.S_rk_list = S_rk_list
        return Linop, rhstot

    def construct_rhstot(self, bc):
        return 0
