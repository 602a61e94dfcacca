"""ULV factorization / solve of HBS matrices in 3-D tensor form, on the MI355X.

Mirror of the module-level API of matAssembly/HBS/ULVsparse_torch.py (SimonDirckx/SslabLU):
compute_ULV(Utens, Dtens, Vtens, Nbvec, device) -> (Qlist, Wlist, Uulist, Rlist, NNvec)
(:231-295) and solve(Umats, Dmats, Qlist, Wlist, Uulist, Rlist, NNvec, rhs, device,
mode='N') (:297-363), where Utens[l] / Vtens[l] are (Nb_l, n_l, k) and Dtens[l] is
(Nb_l, n_l, n_l).  Same device engine as the ULVsparse mirror; the factor lists are
device tensors of this implementation, to be passed back to solve() unchanged.
"""
from __future__ import annotations

from . import ULVsparse


def _flat(X):
    return X.reshape(-1, X.shape[-1])


def compute_ULV(Utens, Dtens, Vtens, Nbvec, device=None):
    fac = Nbvec[0] // Nbvec[1] if len(Nbvec) > 1 and Nbvec[1] else 2
    return ULVsparse.compute_ULV([_flat(u) for u in Utens], [_flat(d) for d in Dtens],
                                 [_flat(v) for v in Vtens], Nbvec, fac=fac)


def solve(Umats, Dmats, Qlist, Wlist, Uulist, Rlist, NNvec, rhs, device=None, mode="N"):
    return ULVsparse.solve(Umats, Dmats, Qlist, Wlist, Uulist, Rlist, NNvec, None, rhs,
                           mode=mode)
