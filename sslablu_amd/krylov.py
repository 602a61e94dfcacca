"""Device GMRES for the interface system, with a direct solver as preconditioner.

The reference closes its 3-D runs with scipy's GMRES on the S operator, preconditioned by
the (compressed) direct solver (cubeTest.py:163-186, on the operator of
multislab/oms.py:267-304).  Here every right-hand-side column runs its own restarted,
right-preconditioned GMRES, all columns advancing together: the vectors live on the GPU
as (n, r) panels, the Gram-Schmidt products and updates are the column-batched kernels
of libsslab_hip.so (sslab_col_dots / sslab_col_axpy / sslab_col_scale), the operator and
preconditioner are callables (typically a block-tridiagonal apply over the library's
GEMM and a BlockTriFactor / HBS / folded solve), and only the small Hessenberg systems
and their Givens rotations are solved on the host, one (k x r) transfer per iteration.

Orthogonalisation is classical Gram-Schmidt applied twice (CGS2), which keeps the basis
orthogonal to working precision with two column-batched passes per step.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_handle

F64 = torch.float64


def _dots(V, k, W, s):
    """(k, r) host array of column-wise products of the first k basis panels with W."""
    n, r = W.shape
    H = torch.empty(k, r, dtype=F64, device=W.device)
    call("sslab_col_dots", n, r, k, ptr(V), V.stride(0), V.stride(1), ptr(W), W.stride(0),
         ptr(H), r, s)
    return H


def _axpy(W, alpha, V, k, H, s):
    n, r = W.shape
    call("sslab_col_axpy", n, r, k, float(alpha), ptr(V), V.stride(0), V.stride(1), ptr(H),
         H.stride(0), ptr(W), W.stride(0), s)


def _scale(W, sc, out, s):
    n, r = W.shape
    call("sslab_col_scale", n, r, ptr(W), W.stride(0), ptr(sc), ptr(out), out.stride(0), s)


def _norms(W, s):
    return np.sqrt(np.maximum(_dots(W[None], 1, W, s).cpu().numpy()[0], 0.0))


@dataclass
class GMRESInfo:
    iterations: int = 0                 # Arnoldi steps (operator applications in cycles)
    restarts: int = 0
    converged: bool = False
    residuals: np.ndarray = None        # final relative residual per column
    history: list = field(default_factory=list)  # max relative residual per step


def gmres(apply_A, B: torch.Tensor, apply_M=None, X0: torch.Tensor | None = None,
          tol: float = 1e-10, restart: int = 30, maxiter: int = 300, stream=None):
    """Solve A X = B column by column (each column its own Krylov space) with right
    preconditioning A M^-1 u = b, x = M^-1 u.  apply_A / apply_M map an (n, r) float64
    device panel to a new one.  Stops when every column's relative residual
    ||b - A x|| / ||b|| (the GMRES estimate, rechecked at each restart) is below tol or
    after maxiter Arnoldi steps.  Returns (X, GMRESInfo)."""
    _lib.require_gpu()
    if B.dim() != 2 or B.dtype != F64 or not B.is_cuda:
        raise ValueError("gmres: B must be a float64 (n, r) device tensor")
    B = B.contiguous()
    n, r = B.shape
    s = stream_handle(stream)
    dev = B.device
    M = apply_M if apply_M is not None else (lambda v: v)
    ones = torch.ones(1, r, dtype=F64, device=dev)
    X = torch.zeros_like(B) if X0 is None else X0.clone().contiguous()
    bnorm = _norms(B, s)
    bnorm = np.where(bnorm > 0, bnorm, 1.0)
    info = GMRESInfo()
    m = max(1, min(restart, n))
    V = torch.empty(m + 1, n, r, dtype=F64, device=dev)
    while True:
        R = B.clone()
        if info.iterations > 0 or X0 is not None:
            _axpy(R, -1.0, apply_A(X)[None], 1, ones, s)      # R = B - A X
        beta = _norms(R, s)
        rel = beta / bnorm
        info.residuals = rel
        if np.all(rel < tol):
            info.converged = True
            return X, info
        if info.iterations >= maxiter:
            return X, info
        inv = torch.tensor(np.where(beta > 0, 1.0 / np.where(beta > 0, beta, 1.0), 0.0),
                           dtype=F64, device=dev)
        _scale(R, inv, V[0], s)
        g = np.zeros((m + 1, r))
        g[0] = beta
        Hh = np.zeros((m + 1, m, r))
        cs = np.zeros((m, r))
        sn = np.zeros((m, r))
        j_done = 0
        for j in range(m):
            w = apply_A(M(V[j])).contiguous()
            h = _dots(V, j + 1, w, s)
            _axpy(w, -1.0, V, j + 1, h, s)
            h2 = _dots(V, j + 1, w, s)
            _axpy(w, -1.0, V, j + 1, h2, s)
            hcol = (h + h2).cpu().numpy()
            hn = _norms(w, s)
            Hh[:j + 1, j] = hcol
            Hh[j + 1, j] = hn
            inv = torch.tensor(np.where(hn > 0, 1.0 / np.where(hn > 0, hn, 1.0), 0.0),
                               dtype=F64, device=dev)
            _scale(w, inv, V[j + 1], s)
            # Givens rotations, column by column (vectorised over the r columns)
            for i in range(j):
                t = cs[i] * Hh[i, j] + sn[i] * Hh[i + 1, j]
                Hh[i + 1, j] = -sn[i] * Hh[i, j] + cs[i] * Hh[i + 1, j]
                Hh[i, j] = t
            den = np.hypot(Hh[j, j], Hh[j + 1, j])
            den_safe = np.where(den > 0, den, 1.0)
            cs[j] = np.where(den > 0, Hh[j, j] / den_safe, 1.0)
            sn[j] = np.where(den > 0, Hh[j + 1, j] / den_safe, 0.0)
            Hh[j, j] = den
            Hh[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            info.iterations += 1
            j_done = j + 1
            est = np.abs(g[j + 1]) / bnorm
            info.history.append(float(est.max()))
            if np.all(est < tol) or info.iterations >= maxiter:
                break
        # y: upper-triangular solve per column, then x += M^-1 (V y)
        k = j_done
        Y = np.zeros((k, r))
        for c in range(r):
            Rk = Hh[:k, :k, c]
            d = np.abs(np.diag(Rk))
            if np.all(d > 0):
                Y[:, c] = np.linalg.solve(np.triu(Rk), g[:k, c])
            else:  # breakdown in this column: least squares on the nonsingular part
                Y[:, c] = np.linalg.lstsq(np.triu(Rk), g[:k, c], rcond=None)[0]
        Yd = torch.tensor(Y, dtype=F64, device=dev)
        U = torch.zeros_like(B)
        _axpy(U, 1.0, V, k, Yd, s)
        _axpy(X, 1.0, M(U).contiguous()[None], 1, ones, s)
        info.restarts += 1


def tridiag_apply(SiM: torch.Tensor, SiP: torch.Tensor, nb: int, m: int, T=None):
    """X -> S X for S = tridiag(SiM, T or I, SiP) with nb blocks of size m (block row i:
    T_i x_i + SiM_i x_{i-1} + SiP_i x_{i+1}), the S-system of multislab/oms.py:267-304.
    SiM / SiP are (nb, m, m) or (1, m, m) (one block shared by every row); the products
    run on the library's batched GEMM."""
    from .hbs import bgemm

    def sel(A, lo, hi):
        return A if A.shape[0] == 1 else A[lo:hi]

    def apply(X):
        X = X.contiguous()
        r = X.shape[1]
        Xb = X.view(nb, m, r)
        if T is None:
            out = X.clone()
        else:
            out = bgemm(T, Xb).reshape(nb * m, r)
        ob = out.view(nb, m, r)
        if nb > 1:
            bgemm(sel(SiM, 1, nb), Xb[:-1], out=ob[1:], beta=1.0)
            bgemm(sel(SiP, 0, nb - 1), Xb[1:], out=ob[:-1], beta=1.0)
        return out

    return apply
