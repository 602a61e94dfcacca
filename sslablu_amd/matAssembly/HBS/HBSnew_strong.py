"""Strong-admissibility HBS compression of a black-box operator, on the MI355X.

Mirror of matAssembly/HBS/HBSnew_strong.py (SimonDirckx/SslabLU): class HBSStrong with
the constructor HBSStrong(A, Nb=None, L=None, device='cpu', tree=None, perm=None,
neighbors=None) (:46-79) -- neighbours from tree.level_adj_list, a callable
neighbors(lvl, t), a per-level list, or by default the 1-D rule |i - j| = 1 on 2^lvl
boxes per level (:42-44, :81-111) --, sample(rk, p=10, del_op=False) with
s = min(maxdeg max(nl, fac rk) + rk + p, N) Gaussian samples (:394-418),
construct(rk, p=10, fast=False, chunk=None) (:420-471) and matvec / rmatvec / @
(:473-525).

Same device engine as the HBStorch_strong mirror (sslablu_amd.hbs_strong); the
reference's fast/chunk options select CPU memory strategies and have no effect.  Every
level uses all s samples (the reference trims level l to g b_l + rk + p of them,
:442-443; more samples only add accuracy).  The test matrices are drawn on the device
from a torch.Generator (seed attribute, default 0); device='cpu' selects the current GPU
like any other value: there is no CPU path.
"""
from __future__ import annotations

import numpy as np
import torch

from ... import _lib
from ...hbs import F64, device_perm
from ...hbs_strong import DeviceHBSStrong
from .HBSnew import _apply_op, fromnumpy


def neighbors_1d_excl(i, nb):
    return [j for j in (i - 1, i + 1) if 0 <= j < nb]


class HBSStrong:
    device_op = True

    def __init__(self, A, Nb=None, L=None, device="cpu", tree=None, perm=None, neighbors=None,
                 seed=0):
        _lib.require_gpu()
        self.A = A
        self.N = A.shape[0]
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.fac = 2
        self.seed = seed
        self._tree = tree
        if tree is not None:
            Nb = tree.nleaves
            L = tree.nlevels - 1
            perm = np.asarray(tree.perm_leaf)
            self._adj = [[set(a) for a in lv] for lv in tree.level_adj_list]
            if getattr(tree, "is_quad", False):
                self.fac = 4
        elif neighbors is None:
            self._adj = [[set(neighbors_1d_excl(t, 2 ** l)) for t in range(2 ** l)]
                         for l in range(L + 1)]
        elif callable(neighbors):
            self._adj = [[{j for j in neighbors(l, t) if j != t} for t in range(2 ** l)]
                         for l in range(L + 1)]
        else:
            self._adj = [[{j for j in lv[t] if j != t} for t in range(len(lv))] for lv in neighbors]
        self.Nb, self.L = Nb, L
        self.nl = self.N // Nb
        self.perm = np.arange(self.N) if perm is None else np.asarray(perm)
        self._h = None
        self.mode = "N"

    def sample(self, rk, p=10, del_op=False):
        if self.N % self.Nb:
            raise ValueError(f"non-uniform boxes: N={self.N} is not divisible by Nb={self.Nb}")
        maxdeg = max((max(len(a) + 1 for a in self._adj[l]) for l in range(2, self.L + 1)),
                     default=1)
        s = min(maxdeg * max(self.nl, self.fac * rk) + rk + p, self.N)
        self.nSamples = s
        dev = self.device
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(self.seed))
        Om = torch.randn(self.N, s, generator=gen, device=dev, dtype=F64)
        Psi = torch.randn(self.N, s, generator=gen, device=dev, dtype=F64)
        pt = device_perm(self.perm, dev)
        self._pt = pt
        rows = (lambda X: X.contiguous()) if pt is None else (
            lambda X: X.index_select(0, pt).contiguous())
        self._Om, self._Psi = rows(Om), rows(Psi)
        self._Y = rows(_apply_op(self.A, Om, False, dev))
        self._Z = rows(_apply_op(self.A, Psi, True, dev))
        if del_op:
            self.A = None
        return self

    def construct(self, rk, p=10, fast=False, chunk=None):
        self._h = DeviceHBSStrong.from_sketches(self._Om, self._Psi, self._Y, self._Z,
                                                self._adj, rk, self.fac, self._pt)
        self._Om = self._Psi = self._Y = self._Z = None
        return self

    def _apply(self, v, transpose=False):
        was_np = not isinstance(v, torch.Tensor)
        x = fromnumpy(v, self.device) if was_np else v.to(self.device, F64)
        y = self._h.apply(x, "T" if transpose else "N")
        return y.cpu().numpy() if was_np else y

    def matvec(self, v):
        return self._apply(v, False)

    def rmatvec(self, v):
        return self._apply(v, True)

    def matmat(self, v):
        return self._apply(v, False)

    def rmatmat(self, v):
        return self._apply(v, True)

    def __matmul__(self, v):
        return self._apply(v, False)

    @property
    def nbytes(self):
        return 0 if self._h is None else self._h.nbytes
