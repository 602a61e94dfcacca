"""Device-resident dense blocks with a numpy-compatible surface.

The reference passes S blocks around as numpy arrays or scipy LinearOperators and its
drivers multiply them with numpy arrays (e.g. ``S_rk_list[i][j] @ E`` in
test/3D/cubeTest.py:150-155).  DeviceMatrix keeps the block in HBM and answers `@`,
.T, matmat / rmatmat / matvec / rmatvec and numpy.asarray: numpy operands give numpy
results, device tensors stay on the device.  Products run on the MFMA GEMM of
libsslab_hip.so.
"""
from __future__ import annotations

import numpy as np
import torch

from ._convert import F64


class DeviceMatrix:
    __array_priority__ = 1000  # make ndarray @ DeviceMatrix defer to __rmatmul__

    def __init__(self, t: torch.Tensor):
        if t.dim() != 2:
            raise ValueError("DeviceMatrix wraps a 2-D tensor")
        self.t = t if t.dtype == F64 else t.to(F64)

    shape = property(lambda self: tuple(self.t.shape))
    dtype = property(lambda self: np.dtype(np.float64))
    ndim = 2
    device_op = True

    def _mul(self, M, x, transpose=False):
        from .ops import dgemm
        was_np = not isinstance(x, torch.Tensor)
        X = torch.as_tensor(np.asarray(x, dtype=np.float64)).to(M.device) if was_np else x.to(M.device, F64)
        one_d = X.dim() == 1
        if one_d:
            X = X[:, None]
        Y = dgemm(M.t() if transpose else M, X)
        if one_d:
            Y = Y[:, 0]
        return Y.cpu().numpy() if was_np else Y

    def __matmul__(self, x):
        if isinstance(x, DeviceMatrix):
            return DeviceMatrix(self._mul(self.t, x.t))
        return self._mul(self.t, x)

    def __rmatmul__(self, x):  # x @ self = (self^T x^T)^T
        was_np = not isinstance(x, torch.Tensor)
        y = self._mul(self.t, np.asarray(x).T if was_np else x.t(), transpose=True)
        return y.T if was_np else y.t()

    def matmat(self, x):
        return self._mul(self.t, x)

    matvec = matmat

    def rmatmat(self, x):
        return self._mul(self.t, x, transpose=True)

    rmatvec = rmatmat

    @property
    def T(self):
        return DeviceMatrix(self.t.t())

    def __array__(self, dtype=None, copy=None):
        a = self.t.detach().cpu().numpy()
        return a.astype(dtype) if dtype is not None else a

    def __neg__(self):
        return DeviceMatrix(-self.t)

    def __repr__(self):
        return f"DeviceMatrix(shape={self.shape}, device={self.t.device})"


def as_tensor_block(b, device):
    """A square block (numpy, torch, DeviceMatrix, or an operator with matmat) as a
    float64 device tensor."""
    if isinstance(b, DeviceMatrix):
        return b.t.to(device)
    if isinstance(b, torch.Tensor):
        return b.to(device=device, dtype=F64)
    if hasattr(b, "matmat") and not isinstance(b, np.ndarray):
        n = b.shape[1]
        eye = torch.eye(n, dtype=F64, device=device)
        y = b.matmat(eye) if getattr(b, "device_op", False) or hasattr(b, "_h") else b.matmat(np.eye(n))
        return torch.as_tensor(np.asarray(y) if not isinstance(y, torch.Tensor) else y).to(device, F64)
    return torch.as_tensor(np.asarray(b, dtype=np.float64)).to(device)
