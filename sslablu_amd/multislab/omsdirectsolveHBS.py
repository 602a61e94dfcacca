"""Older identity-diagonal HBS Thomas API, on the MI355X.

Mirror of multislab/omsdirectsolveHBS.py (SimonDirckx/SslabLU): Sprime_Linop(Sl,
Sprime_prev, Sr, id=False) (:19-47), build_block_tridiagonal_solver(S_rk_list, tree,
quad, rk) -> (Sl, Sprime, Sr) (:53-98) and block_tridiagonal_solve(OMS, T, rhs)
(:100-138).  S'_1 = I - S_{1,0} S_{0,1} and S'_{i+1} = I - S_{i+1,i} S'_i^{-1} S_{i,i+1}
are device operators, each HBS-compressed on `tree` with rank `rk` and ULV-factored
(sslablu_amd.matAssembly.HBS.HBSnew).  Right-hand sides: (N,) or (N, r), numpy or
float64 device tensors; the output keeps the input's kind and ndim.
"""
from __future__ import annotations

import numpy as np
import torch

from ..direct_solve.omsdirectsolveHBS import Sprime_Linop, _back, _mm, _sv, _t, id_op
from ..matAssembly.HBS import HBSnew

__all__ = ["Sprime_Linop", "build_block_tridiagonal_solver", "block_tridiagonal_solve"]


def build_block_tridiagonal_solver(S_rk_list, tree, quad, rk):
    m = S_rk_list[0][0].shape[0]
    n = len(S_rk_list) - 1
    Sl = [S_rk_list[k][0] for k in range(1, n + 1)]
    Sr = [S_rk_list[k][-1] for k in range(n)]
    Sprime = [id_op(m)]
    for i in range(1, n + 1):
        op = Sprime_Linop(Sl[i - 1], Sprime[i - 1], Sr[i - 1], id=(i == 1))
        h = HBSnew.HBSMAT(op, tree, quad)
        h.construct(rk, compute_ULV=True)
        Sprime.append(h)
    return Sl, Sprime, Sr


def block_tridiagonal_solve(OMS, T, rhs):
    Sl, Sp, Sr = T
    n = len(Sl)
    idx = [np.asarray(list(r)) for r in OMS.glob_target_dofs]
    D, was_np, one_d = _t(rhs)
    d = D.clone()
    for i in range(1, n + 1):
        v = d[idx[i - 1]] if i == 1 else _sv(Sp[i - 1], d[idx[i - 1]])
        d[idx[i]] = d[idx[i]] - _mm(Sl[i - 1], v)
    x = torch.zeros_like(d)
    x[idx[n]] = _sv(Sp[n], d[idx[n]])
    for i in range(n - 1, 0, -1):
        x[idx[i]] = _sv(Sp[i], d[idx[i]] - _mm(Sr[i], x[idx[i + 1]]))
    x[idx[0]] = d[idx[0]] - _mm(Sr[0], x[idx[1]])  # S'_0 = I
    return _back(x, was_np, one_d)
