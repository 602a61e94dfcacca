"""One red-black reduction level on explicit block lists (GPU primitives).

Mirror of RedBlackSolver.build_block_RB_solver_level
(direct_solve/omsdirectsolve.py:283-333) for callers that drive the levels
themselves; the solver classes use the fused level kernels of libsslab_hip.so.
"""
from __future__ import annotations

import torch

from .._convert import F64
from .._lib import call, ptr, stream_handle
from ..ops import dgemm


def _dev(x, device):
    return torch.as_tensor(x).to(device=device, dtype=F64).contiguous()


def lu_factor_gpu(A: torch.Tensor):
    """(lu, piv) of a square device matrix (scipy.linalg.lu_factor semantics)."""
    n = A.shape[0]
    lu = A.to(dtype=F64).contiguous().clone()
    piv = torch.empty(n, dtype=torch.int32, device=lu.device)
    info = torch.zeros(1, dtype=torch.int32, device=lu.device)
    call("sslab_dgetrf_batched", n, ptr(lu), n, 0, ptr(piv), 0, ptr(info), 1, stream_handle())
    return lu, piv


def lu_solve_gpu(lu_piv, B: torch.Tensor) -> torch.Tensor:
    """A^{-1} B from an (lu, piv) pair (scipy.linalg.lu_solve semantics)."""
    lu, piv = lu_piv
    dev = lu.device if isinstance(lu, torch.Tensor) else B.device
    lu = _dev(lu, dev)
    piv = torch.as_tensor(piv).to(device=dev, dtype=torch.int32).contiguous()
    n = lu.shape[0]
    Bm = _dev(B, dev)
    one_d = Bm.dim() == 1
    if one_d:
        Bm = Bm[:, None]
    r = Bm.shape[1]
    X = torch.empty_like(Bm)
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    call("sslab_dgetrs_batched", n, r, ptr(lu), n, 0, ptr(piv), 0, ptr(Bm), r, 0, ptr(X), r, 0,
         ptr(perm), 1, stream_handle())
    return X[:, 0] if one_d else X


def rb_level_gpu(m, nSlabs, RB_level, cyclic, device):
    SiM = [_dev(s, device) for s in RB_level[0]]
    T = [_dev(t, device) for t in RB_level[1]]
    T_lu = [(_dev(l, device), torch.as_tensor(p).to(device=device, dtype=torch.int32))
            for (l, p) in RB_level[2]]
    SiP = [_dev(s, device) for s in RB_level[3]]
    Bn, An, Cn = [], [], []
    for i in range(0, nSlabs, 2):
        B = T[i].clone()
        if i > 0 or cyclic:
            dgemm(SiM[i], lu_solve_gpu(T_lu[i - 1], SiP[(i - 1) % nSlabs]), B, -1.0, 1.0)
        if i < nSlabs - 1 or cyclic:
            dgemm(SiP[i], lu_solve_gpu(T_lu[(i + 1) % nSlabs], SiM[(i + 1) % nSlabs]), B, -1.0, 1.0)
        Bn.append(B)
        An.append(dgemm(SiM[i], lu_solve_gpu(T_lu[i - 1], SiM[(i - 1) % nSlabs]), alpha=-1.0))
        Cn.append(dgemm(SiP[i], lu_solve_gpu(T_lu[(i + 1) % nSlabs], SiP[(i + 1) % nSlabs]), alpha=-1.0))
    if len(Bn) == 1 and cyclic:
        Bn[0] = Bn[0] + An[0] + Cn[0]
    return (An, Bn, [lu_factor_gpu(b) for b in Bn], Cn)
