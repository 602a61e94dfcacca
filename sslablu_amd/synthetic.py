"""Synthetic block-tridiagonal systems generated on the device.

``gaussian_blocktri``: nb diagonal blocks of size n from 1-D point sets on [0, 1]
(each block's points sorted U(0,1), seeded), Gaussian-kernel entries
K(x, y) = exp(-|x - y|^2 / h), coupling blocks scaled by `coupling` and the diagonal
shifted by n for dominance:
    T_b = K(x_b, x_b) + n I,  SiM_b = coupling K(x_b, x_{b-1}),  SiP_b = coupling K(x_b, x_{b+1}).
The point sets come from numpy.random.default_rng(seed) (so a CPU check can rebuild
the same blocks); the kernel matrices are formed on the GPU.
"""
from __future__ import annotations

import numpy as np
import torch

F64 = torch.float64


def gaussian_points(nb: int, n: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.sort(rng.random((nb, n)), axis=1)


def gaussian_blocktri(nb: int, n: int, seed: int = 0, h: float = 0.01, coupling: float = 0.1,
                      device=None):
    """(T, SiM, SiP) as (nb, n, n) float64 device tensors; SiM[0] and SiP[nb-1] are zero."""
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    x = torch.tensor(gaussian_points(nb, n, seed), dtype=F64, device=dev)

    def K(a, b):
        return torch.exp(-(a[:, :, None] - b[:, None, :]) ** 2 / h)

    T = K(x, x)
    T.diagonal(dim1=1, dim2=2).add_(float(n))
    SiM = torch.zeros_like(T)
    SiP = torch.zeros_like(T)
    if nb > 1:
        SiM[1:] = coupling * K(x[1:], x[:-1])
        SiP[:-1] = coupling * K(x[:-1], x[1:])
    return T, SiM, SiP
