"""Dense block-tridiagonal direct solvers of the S-system, on the MI355X.

Drop-in for direct_solve/omsdirectsolve.py (SimonDirckx/SslabLU): the same class
names, constructor arguments, method signatures and attributes, backed by
libsslab_hip.so.  Differences (all supersets):

* right-hand sides may be (N,) or (N, r) for every solver (the reference's
  RedBlackSolver takes (N,) only, omsdirectsolve.py:221-279), and numpy inputs give
  numpy outputs while torch device tensors stay on the device;
* RedBlackSolver accepts any block count (the reference requires a power of two,
  omsdirectsolve.py:200-202) and periodic systems use a Woodbury correction;
* factors live in HBM; A/B/C/RB expose device tensors instead of numpy arrays.
"""
from __future__ import annotations

from abc import ABC, abstractmethod

import torch

from .. import _lib
from .._convert import F64, blocks_to_device, from_device, index_list, to_device
from ..blocktri import BlockTriFactor
from ..ops import dgemm


class DirectSolver(ABC):
    """Abstract direct solver (omsdirectsolve.py:13-27)."""

    def __init__(self, m, cyclic=False):
        self.m = m
        self.cyclic = cyclic

    @abstractmethod
    def factorize(self, S_rk_list, T):
        """Factor the block system given its off-diagonal pairs and diagonal blocks."""

    @abstractmethod
    def solve(self, rhs):
        """Solve with the stored factorization."""


def reconstruct_from_lu_factor(Tinv):
    """Dense T from an (lu, piv) pair with scipy lu_factor semantics
    (omsdirectsolve.py:31-46): T = P^T L U.  Works for numpy or torch pairs."""
    lu, piv = Tinv
    is_np = not isinstance(lu, torch.Tensor)
    lu_t = torch.as_tensor(lu, dtype=F64)
    piv_t = torch.as_tensor(piv).to("cpu").long()
    n = lu_t.shape[0]
    eye = torch.eye(n, dtype=F64, device=lu_t.device)
    M = (torch.tril(lu_t, -1) + eye) @ torch.triu(lu_t)
    for i in range(n - 1, -1, -1):
        p = int(piv_t[i])
        if p != i:
            M[[i, p]] = M[[p, i]]
    return M.cpu().numpy() if is_np else M


def _split_system(S_rk_list, T, m, device):
    """Device stacks (T, SiM, SiP) and the corner blocks (E, F) of S_rk_list, with the
    reference's indexing: SiM_i = S_rk_list[i][0], SiP_i = S_rk_list[i][-1]
    (omsdirectsolve.py:85,89,197-198)."""
    nb = len(S_rk_list)
    if T is None:
        Td = torch.eye(m, dtype=F64, device=device).expand(nb, m, m).contiguous()
    else:
        Td = blocks_to_device(T, device)
    SiM = blocks_to_device([s[0] for s in S_rk_list], device)
    SiP = blocks_to_device([s[-1] for s in S_rk_list], device)
    return Td, SiM, SiP, SiM[0], SiP[nb - 1]


class _GPUSolverBase(DirectSolver):
    _method = "rb"

    def __init__(self, m, cyclic=False):
        super().__init__(m, cyclic)
        _lib.require_gpu()
        self.device = torch.device("cuda", torch.cuda.current_device())
        self._fac = None

    def _factorize(self, S_rk_list, T):
        m = S_rk_list[0][0].shape[0]
        self.m = m
        Td, SiM, SiP, E, F = _split_system(S_rk_list, T, m, self.device)
        fac = BlockTriFactor(self._method, self.device)
        if self.cyclic:
            fac.factorize(Td, SiM, SiP, E=E, F=F)
        else:
            fac.factorize(Td, SiM, SiP)
        self._fac = fac
        self.nSlabs = len(S_rk_list)
        return fac

    def _solve(self, rhs, glob_target_dofs=None):
        if self._fac is None:
            raise ValueError("Factorization not set")
        fac = self._fac
        V, was_np, one_d = to_device(rhs, self.device)
        idx = index_list(glob_target_dofs, fac.nb, fac.m)
        if idx is not None:
            idx = idx.to(self.device)
            V = V.index_select(0, idx)
        V = V.contiguous().clone()
        X = torch.empty_like(V)
        fac.solve_(V, X)
        if idx is not None:
            out = torch.zeros((max(int(idx.max()) + 1, fac.nb * fac.m), X.shape[1]), dtype=F64,
                              device=self.device)
            out[idx] = X
            X = out
        return from_device(X, was_np, one_d)


class BlockTridiagonalSolver(_GPUSolverBase):
    """Thomas elimination with a general diagonal T and an optional periodic
    Woodbury correction (omsdirectsolve.py:48-175), on the GPU."""

    _method = "thomas"

    def factorize_helper(self, S_rk_list, T):
        """Forward-elimination factors A_i, B_i, C_i (omsdirectsolve.py:50-100)."""
        self._factorize(S_rk_list, T)

    def solve_helper(self, rhs, glob_target_dofs=None):
        """Thomas sweeps without the periodic correction (omsdirectsolve.py:102-141)."""
        cyc = self._fac.cyclic
        self._fac.cyclic = False
        try:
            return self._solve(rhs, glob_target_dofs)
        finally:
            self._fac.cyclic = cyc

    def factorize(self, S_rk_list, T, glob_target_dofs=None):
        self._factorize(S_rk_list, T)

    def solve(self, rhs, glob_target_dofs=None):
        return self._solve(rhs, glob_target_dofs)

    # reference attributes --------------------------------------------------------
    def _views(self):
        fac = self._fac
        nb, m = fac.nb, fac.m
        f = fac.factor
        m2 = m * m
        A = f[: (nb - 1) * m2].view(max(nb - 1, 0), m, m)
        LU = f[(2 * nb - 1) * m2:(3 * nb - 1) * m2].view(nb, m, m)
        C = f[(3 * nb - 1) * m2:(4 * nb - 2) * m2].view(max(nb - 1, 0), m, m)
        return A, LU, C

    @property
    def A(self):
        """A_i = S_{i,i-1} B_{i-1}^{-1}, i = 1..nb-1 (device tensors)."""
        return list(self._views()[0])

    @property
    def B(self):
        """(lu, piv) pairs of the eliminated diagonal blocks B_i (device tensors)."""
        _, LU, _ = self._views()
        return [(LU[i], self._fac.ipiv[i]) for i in range(self._fac.nb)]

    @property
    def C(self):
        """C_i = S_rk_list[i][-1], i = 0..nb-2 (device tensors)."""
        return list(self._views()[2])

    @property
    def smw_block(self):
        """Dense N x N Woodbury block Z (I + V^T Z)^{-1} V^T of a periodic system
        (omsdirectsolve.py:166); None otherwise.  Built on demand: the solver itself
        applies the correction in factored form."""
        fac = self._fac
        if fac is None or not fac.cyclic:
            return None
        nb, m = fac.nb, fac.m
        ZK = dgemm(fac.Z, fac.K)  # N x 2m
        out = torch.zeros(nb * m, nb * m, dtype=F64, device=self.device)
        out[:, (nb - 1) * m:] = ZK[:, :m]
        out[:, :m] += ZK[:, m:]
        return out


class RedBlackSolver(_GPUSolverBase):
    """Red-black cyclic reduction (omsdirectsolve.py:177-333), on the GPU: every
    level eliminates the odd blocks with batched MFMA kernels."""

    _method = "rb"

    def factorize(self, S_rk_list, T):
        self._factorize(S_rk_list, T)

    def solve(self, rhs):
        return self._solve(rhs)

    @property
    def RB(self):
        """Per-level device factors: dicts with the odd-node blocks
        TG = [T^{-1} | T^{-1} SiM | T^{-1} SiP] and the even-node blocks
        WW = [SiM T_{i-1}^{-1} | SiP T_{i+1}^{-1}] (the reference keeps
        (SiM, T, lu(T), SiP) per level instead; the solution is the same)."""
        fac = self._fac
        if fac is None:
            return []
        nb, m = fac.nb, fac.m
        m2 = m * m
        out, off, n = [], 0, nb
        while n > 1:
            nE, nO = (n + 1) // 2, n // 2
            TG = fac.factor[off:off + nO * 3 * m2].view(nO, m, 3 * m)
            off += nO * 3 * m2
            WW = fac.factor[off:off + nE * 2 * m2].view(nE, m, 2 * m)
            off += nE * 2 * m2
            out.append({"n": n, "TG": TG, "WW": WW})
            n = nE
        out.append({"n": 1, "Tinv": fac.factor[off:off + m2].view(m, m)})
        return out

    def build_block_RB_solver_level(self, m, nSlabs, RB_level, cyclic=False):
        """One reduction level from (SiM, T, T_lu, SiP) block lists
        (omsdirectsolve.py:283-333): returns (A, B, lu(B), C) for the even blocks.
        Computed on the GPU with the batched LU / GEMM primitives."""
        from ._level import rb_level_gpu
        return rb_level_gpu(m, nSlabs, RB_level, cyclic, self.device)
