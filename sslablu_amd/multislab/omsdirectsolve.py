"""Identity-diagonal block-tridiagonal solvers of the oms S-system (older API), on the GPU.

Mirror of multislab/omsdirectsolve.py (SimonDirckx/SslabLU):
build_block_tridiagonal_solver(S_rk_list) -> (A, B, C) (:17-68),
block_tridiagonal_solve(OMS, T, rhs) (:70-99),
build_block_cyclic_tridiagonal_solver(OMS, S_rk_list, rhs_list, Ntot, nc)
-> ((A, B, C), smw) (:102-145) and block_cyclic_tridiagonal_solve(OMS, T, smw, d)
(:147-160).  The system is [I S_12; S_21 I S_23; ...] (plus corner blocks E, F when
periodic); the factors live in HBM (Thomas with B_0 = I), A/B/C are device views, and
`smw` applies the Woodbury correction in factored form (the reference forms the dense
Ntot x Ntot matrix, :139).
"""
from __future__ import annotations

import numpy as np
import torch

from .._convert import F64, blocks_to_device, from_device, index_list, to_device
from ..blocktri import BlockTriFactor


class _Factors(tuple):
    """(A, B, C) device views, carrying the GPU factorization in .fac."""

    fac = None


def _factor(S_rk_list, cyclic):
    dev = torch.device("cuda", torch.cuda.current_device())
    nb = len(S_rk_list)
    m = S_rk_list[0][0].shape[0]
    SiM = blocks_to_device([s[0] for s in S_rk_list], dev)
    SiP = blocks_to_device([s[-1] for s in S_rk_list], dev)
    eye = torch.eye(m, dtype=F64, device=dev).expand(nb, m, m).contiguous()
    fac = BlockTriFactor("thomas", dev)
    if cyclic:
        fac.factorize(eye, SiM, SiP, E=SiM[0], F=SiP[nb - 1])
    else:
        fac.factorize(eye, SiM, SiP)
    f = fac.factor
    m2 = m * m
    A = list(f[: (nb - 1) * m2].view(max(nb - 1, 0), m, m))
    LU = f[(2 * nb - 1) * m2:(3 * nb - 1) * m2].view(nb, m, m)
    B = [(LU[i], fac.ipiv[i]) for i in range(nb)]
    C = list(f[(3 * nb - 1) * m2:(4 * nb - 2) * m2].view(max(nb - 1, 0), m, m))
    out = _Factors((A, B, C))
    out.fac = fac
    return out


def build_block_tridiagonal_solver(S_rk_list):
    """Thomas factors of [I S_12; S_21 I ...] (multislab/omsdirectsolve.py:17-68)."""
    return _factor(S_rk_list, cyclic=False)


def _solve(fac, rhs, indices, cyclic):
    V, was_np, one_d = to_device(rhs, fac.device)
    idx = index_list(indices, fac.nb, fac.m) if indices is not None else None
    if idx is not None:
        V = V.index_select(0, idx.to(fac.device))
    V = V.contiguous().clone()
    X = torch.empty_like(V)
    keep = fac.cyclic
    fac.cyclic = cyclic and keep
    try:
        fac.solve_(V, X)
    finally:
        fac.cyclic = keep
    if idx is not None:
        out = torch.zeros_like(X)
        out[idx.to(fac.device)] = X
        X = out
    return from_device(X, was_np, one_d)


def block_tridiagonal_solve(OMS, T, rhs):
    """Solve with factors from build_block_tridiagonal_solver (:70-99)."""
    indices = getattr(OMS, "glob_target_dofs", None) if OMS is not None else None
    return _solve(T.fac, rhs, indices or None, cyclic=False)


class _Woodbury:
    """smw such that x = y - smw @ y applies (I + U V^T T^{-1})^{-1}-type correction
    of the periodic corner blocks: smw @ y = Z (I + V^T Z)^{-1} V^T y."""

    def __init__(self, fac):
        self.fac = fac

    def __matmul__(self, y):
        fac = self.fac
        Y, was_np, one_d = to_device(y, fac.device)
        m, nb = fac.m, fac.nb
        VtY = torch.cat([Y[(nb - 1) * m:], Y[:m]], dim=0)
        from ..ops import dgemm
        t = dgemm(fac.K, VtY.contiguous())
        out = dgemm(fac.Z, t)
        return from_device(out, was_np, one_d)


def build_block_cyclic_tridiagonal_solver(OMS, S_rk_list, rhs_list, Ntot, nc):
    """Factors of the periodic system with corners E = S_rk_list[0][0] and
    F = S_rk_list[-1][1] (:102-145): ((A, B, C), smw)."""
    T = _factor(S_rk_list, cyclic=True)
    return T, _Woodbury(T.fac)


def block_cyclic_tridiagonal_solve(OMS, T, smw_block, d):
    """y = T \\ d, x = y - smw_block @ y (:147-160)."""
    indices = getattr(OMS, "glob_target_dofs", None) if OMS is not None else None
    y = _solve(T.fac, d, indices or None, cyclic=False)
    return y - smw_block @ y
