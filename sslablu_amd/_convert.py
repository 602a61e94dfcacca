"""numpy <-> device conversions shared by the reference-API mirrors.

The reference API takes numpy float64 arrays and returns numpy arrays of the same
ndim; the mirrors additionally accept torch tensors (kept on the device, returned on
the device) so callers can stay resident in HBM.
"""
from __future__ import annotations

import numpy as np
import torch

F64 = torch.float64


def to_device(x, device):
    """(tensor on `device`, was_numpy, was_1d) for an array-like right-hand side."""
    was_numpy = not isinstance(x, torch.Tensor)
    t = torch.as_tensor(np.asarray(x, dtype=np.float64)) if was_numpy else x
    one_d = t.dim() == 1
    t = t.to(device=device, dtype=F64)
    if one_d:
        t = t[:, None]
    return t, was_numpy, one_d


def from_device(t: torch.Tensor, was_numpy: bool, one_d: bool):
    if one_d:
        t = t[:, 0]
    if was_numpy:
        return t.detach().cpu().numpy()
    return t


def blocks_to_device(blocks, device):
    """list of m x m blocks (numpy, torch, DeviceMatrix or HBS operators) -> (n, m, m)
    contiguous float64 device tensor."""
    if isinstance(blocks, torch.Tensor):
        return blocks.to(device=device, dtype=F64).contiguous()
    from .linop import as_tensor_block
    return torch.stack([as_tensor_block(b, device) for b in blocks]).contiguous()


def index_list(glob_target_dofs, nb, m):
    """Concatenated global dof indices of the blocks, or None when the blocks are the
    contiguous ranges [i*m, (i+1)*m) in order (the reference's default layout,
    multislab/oms.py:177)."""
    if glob_target_dofs is None:
        return None
    idx = np.concatenate([np.asarray(list(r), dtype=np.int64) for r in glob_target_dofs])
    if idx.shape[0] == nb * m and np.array_equal(idx, np.arange(nb * m)):
        return None
    return torch.as_tensor(idx)
