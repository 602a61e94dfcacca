"""Thin device-tensor wrappers over the dense primitives of libsslab_hip.so."""
from __future__ import annotations

import torch

from ._convert import F64
from ._lib import call, ptr, stream_handle


def _gemm(name, dtype, A, B, C, alpha, beta, stream):
    def op(X):
        if X.stride(1) == 1:
            return X, 0, X.stride(0)
        if X.stride(0) == 1:
            return X, 1, X.stride(1)
        X = X.contiguous()
        return X, 0, X.stride(0)

    if A.dtype != dtype or B.dtype != dtype:
        raise TypeError(f"{name}: {dtype} operands required")
    M, K = A.shape
    K2, N = B.shape
    if K != K2:
        raise ValueError(f"{name}: inner dimensions differ")
    A, ta, lda = op(A)
    B, tb, ldb = op(B)
    if C is None:
        C = torch.empty(M, N, dtype=dtype, device=A.device)
        beta = 0.0
    if C.dtype != dtype or C.stride(1) != 1:
        raise ValueError(f"{name}: C needs dtype {dtype} and unit column stride")
    call(f"sslab_{name}_strided_batched", ta, tb, M, N, K, float(alpha), ptr(A), lda, 0, ptr(B),
         ldb, 0, float(beta), ptr(C), C.stride(0), 0, 1, stream_handle(stream))
    return C


def dgemm(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor | None = None, alpha: float = 1.0,
          beta: float = 0.0, stream=None) -> torch.Tensor:
    """C = alpha A B + beta C for 2-D float64 device tensors with unit column stride
    (a transposed view with unit row stride is passed as op(X) = X^T)."""
    return _gemm("dgemm", F64, A, B, C, alpha, beta, stream)


def sgemm(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor | None = None, alpha: float = 1.0,
          beta: float = 0.0, stream=None) -> torch.Tensor:
    """FP32 version of dgemm (FP32 MFMA, FP32 accumulation; the inner dimension must be
    positive)."""
    return _gemm("sgemm", torch.float32, A, B, C, alpha, beta, stream)
