"""Strong-admissibility HBS matrices from caller-supplied sketches, on the MI355X.

Mirror of matAssembly/HBS/HBStorch_strong.py (SimonDirckx/SslabLU): class HBSMAT with the
constructor HBSMAT(A=None, device=None, tree=None, quad=False) (:925-975),
construct(rk, Om, Psi, Y, Z, group_chunk=None, store_device=None, empty_cache=True,
fast=None, profile=True, profile_detail=True) (:1000-1031) -- test matrices and sketches
Y = A Om, Z = A^T Psi in the original point order, permuted to leaf order here --,
nbytes (:993-998), to_device (:1109-1119), print_profile (:1121-1203), T, matvec,
rmatvec, matmat, rmatmat, @ (:1205-1303) and tree (:1305-1310); the near-field
neighbours of every level come from tree.level_adj_list.

The numerics run in sslablu_amd.hbs_strong.DeviceHBSStrong (batched MFMA GEMMs, LU and
pivoted-QR kernels of libsslab_hip.so).  The factors stay in HBM from the start: the
reference parks finished levels in host memory during construction and moves them back
afterwards (store_device, :953-966); group_chunk, store_device and empty_cache are
accepted and have no effect.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from ... import _lib
from ...hbs import F64, device_perm
from ...hbs_strong import DeviceHBSStrong
from .HBSnew import fromnumpy


class HBSMAT:
    """Strong-admissibility HBS approximation of a square map on a cluster tree."""

    device_op = True

    def __init__(self, A=None, device=None, tree=None, quad=False):
        _lib.require_gpu()
        self.device = (torch.device(device) if device not in (None, "cpu", "cuda")
                       else torch.device("cuda", torch.cuda.current_device()))
        self.mode = "N"
        self._h = None
        self._tree = None
        self.dtype = np.float64
        self.quad = quad
        self.fac = 4 if quad else 2
        self.tSample = 0.0
        self.tConstruct = 0.0
        self.tCompress = 0.0
        self.nSamples = 0
        self.Nbvec = []
        self.store_device = self.device
        self.group_chunk = None
        if A is not None:
            self.A = A
            self.shape = tuple(A.shape)
        if tree is not None:
            self.tree = tree
            self.perm = np.asarray(tree.perm_leaf)
            self.Nb = tree.nleaves
            self.shape = (len(self.perm), len(self.perm))
            self.nl = self.shape[0] // self.Nb
            self.L = tree.nlevels

    @property
    def tree(self):
        return self._tree

    @tree.setter
    def tree(self, t):
        self._tree = t
        if t is not None and getattr(t, "is_quad", False):
            self.quad, self.fac = True, 4

    # -- construction (HBStorch_strong.py:1000-1107) ------------------------------------
    def construct(self, rk, Om, Psi, Y, Z, group_chunk=None, store_device=None,
                  empty_cache=True, fast=None, profile=True, profile_detail=True):
        if self._tree is None or self._tree.level_adj_list is None:
            raise ValueError("HBSMAT.construct: a cluster tree with adjacency lists is needed")
        t0 = time.perf_counter()
        dev = self.device
        pt = device_perm(self.perm, dev)

        def rows(X):
            X = fromnumpy(X, dev) if not isinstance(X, torch.Tensor) else X.to(dev, F64)
            return X.contiguous() if pt is None else X.index_select(0, pt).contiguous()

        Om, Psi, Y, Z = (rows(X) for X in (Om, Psi, Y, Z))
        self.nSamples = Om.shape[1]
        self._h = DeviceHBSStrong.from_sketches(Om, Psi, Y, Z, self._tree.level_adj_list, rk,
                                                self.fac, pt)
        torch.cuda.synchronize(dev)
        self.tCompress = time.perf_counter() - t0
        self.tConstruct = self.tCompress
        self.Nbvec = [self.Nb] + [lv["nb"] // self.fac for lv in self._h.levels]
        return self

    def to_device(self, device=None):
        """The factors are device-resident already; kept for the reference's interface."""
        return self

    def print_profile(self, sort_by=None):
        print(f"HBSMAT (strong) construct: {self.tConstruct:.4f} s, {self.nSamples} samples, "
              f"{len(self.Umats)} compressed levels, {self.nbytes / 1e6:.2f} MB")

    # -- factors (lists, leaf level first) ---------------------------------------------
    @property
    def Umats(self):
        return [] if self._h is None else [lv["U"] for lv in self._h.levels]

    @property
    def Vmats(self):
        return [] if self._h is None else [lv["V"] for lv in self._h.levels]

    @property
    def Emats(self):
        """Near-field tiles per level as (nb, b, g b) rows (slot order: sorted near
        field), then the dense base."""
        return [] if self._h is None else [lv["E"] for lv in self._h.levels] + [self._h.B]

    @property
    def nbytes(self):
        return 0 if self._h is None else self._h.nbytes

    # -- apply (HBStorch_strong.py:1205-1303) -------------------------------------------
    @property
    def T(self):
        view = object.__new__(self.__class__)
        view.__dict__ = self.__dict__.copy()
        view.mode = "T" if self.mode == "N" else "N"
        return view

    def _wrap(self, v, mode):
        was_np = not isinstance(v, torch.Tensor)
        x = fromnumpy(v, self.device) if was_np else v.to(self.device, F64)
        y = self._h.apply(x, mode)
        return y.cpu().numpy() if was_np else y

    def matmat(self, v):
        return self._wrap(v, self.mode)

    def rmatmat(self, v):
        return self._wrap(v, "T" if self.mode == "N" else "N")

    def matvec(self, v):
        return self.matmat(v)

    def rmatvec(self, v):
        return self.rmatmat(v)

    def __matmul__(self, v):
        return self.matmat(v)
