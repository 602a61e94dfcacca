"""ULV factorization / solve of HBS matrices in reduced form, on the MI355X.

Mirror of the module-level API of matAssembly/HBS/ULVsparse.py (SimonDirckx/SslabLU):
compute_ULV(Umats, Dmats, Vmats, Nbvec) -> (Qlist, Wlist, Uulist, Rlist, NNvec) and
solve(Umats, Dmats, Qlist, Wlist, Uulist, Rlist, NNvec, Nbvec, rhs, mode='N')
(ULVsparse.py:231-358).  The factor lists are device tensors of this implementation
(Rlist holds the inverted triangular blocks R11^{-1}, the root entry holds
R^{-1} Q^T); pass them back to solve() unchanged.

The device factorization travels with the returned lists (an attribute of the tensor
Qlist[0]), so it lives exactly as long as the caller keeps the factors, and solve()
refuses lists that compute_ULV did not return.
"""
from __future__ import annotations

import numpy as np
import torch

from ...hbs import DeviceHBS, F64

_ATTR = "_sslab_ulv_factor"


def _blocks(X, nb, dev):
    X = torch.as_tensor(np.asarray(X, dtype=np.float64)).to(dev) if not isinstance(X, torch.Tensor) else X.to(dev, F64)
    return X.reshape(nb, -1, X.shape[1]).contiguous()


def _device():
    return torch.device("cuda", torch.cuda.current_device())


def compute_ULV(Umats, Dmats, Vmats, Nbvec, fac=None):
    dev = _device()
    if fac is None:
        fac = Nbvec[0] // Nbvec[1] if len(Nbvec) > 1 and Nbvec[1] else 2
    U = [_blocks(u, Nbvec[i], dev) for i, u in enumerate(Umats)]
    V = [_blocks(v, Nbvec[i], dev) for i, v in enumerate(Vmats)]
    D = [_blocks(d, Nbvec[i], dev) for i, d in enumerate(Dmats)]
    h = DeviceHBS(U, D, V, fac)
    lev = h.ulv_factor("N")
    Qlist = [l["Q"].reshape(-1, l["Q"].shape[2]) for l in lev if not l["root"]] + [lev[-1]["G"][0]]
    Wlist = [l["W"].reshape(-1, l["W"].shape[2]) for l in lev if not l["root"]]
    Uulist = [l["Uu"] for l in lev if not l["root"]]
    Rlist = [l["R11inv"] for l in lev if not l["root"]] + [lev[-1]["G"][0]]
    NNvec = np.concatenate([[0], np.cumsum([(l["n"] - l["k"]) * l["nb"] for l in lev if not l["root"]]
                                           + [lev[-1]["G"].shape[1]])]).astype(np.int64)
    setattr(Qlist[0], _ATTR, h)
    return Qlist, Wlist, Uulist, Rlist, NNvec


def solve(Umats, Dmats, Qlist, Wlist, Uulist, Rlist, NNvec, Nbvec, rhs, mode="N"):
    h = getattr(Qlist[0], _ATTR, None) if len(Qlist) else None
    if h is None:
        raise ValueError("ULV factors must come from this module's compute_ULV")
    was_np = not isinstance(rhs, torch.Tensor)
    b = torch.as_tensor(np.asarray(rhs, dtype=np.float64)).to(h.device) if was_np else rhs
    x = h.solve(b, mode)
    return x.cpu().numpy() if was_np else x
