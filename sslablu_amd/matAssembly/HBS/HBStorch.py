"""HBS matrices from caller-supplied sketches, on the MI355X.

Mirror of matAssembly/HBS/HBStorch.py (SimonDirckx/SslabLU): class HBSMAT with the
constructor HBSMAT(A=None, device=None, tree=None, quad=False) (:102-140),
construct(rk, Om, Psi, Y, Z, compute_ULV=False, fast=False) (:163-167),
constructHBS(rk, Om0, Psi0, Y0, Z0, fast=False) (:169-241) -- the test matrices and
sketches Y = A Om, Z = A^T Psi come from the caller in the original (unpermuted) order
and are permuted to leaf order here --, constructHBS_ULV(rk) (:242-341), matmat /
rmatmat / matvec / rmatvec / @ / T (:343-457), tree (:459-465), compute_ULV (:467) and
solve(b, device, mode='N') (:470-493).

The numerics are the device engine shared with the HBSnew mirror
(sslablu_amd.hbs.DeviceHBS).  Difference, on purpose: the leaf D blocks stay in HBM (the
reference pins them in host memory, :236-239); 288 GB of HBM holds them.
"""
from __future__ import annotations

import torch

from ...hbs import DeviceHBS, device_perm
from .HBSnew import HBSMAT as _HBSnewMAT
from .HBSnew import fromnumpy


class HBSMAT(_HBSnewMAT):
    """HBS approximation built from given sketches (HBStorch.py:85-493)."""

    def __init__(self, A=None, device=None, tree=None, quad=False):
        super().__init__(A=A, tree=tree, quad=quad, device=device)

    def construct(self, rk, Om=None, Psi=None, Y=None, Z=None, compute_ULV=False, fast=False):
        if compute_ULV:
            self.constructHBS_ULV(rk, fast=fast)
        else:
            self.constructHBS(rk, Om, Psi, Y, Z, fast=fast)

    def constructHBS(self, rk, Om0, Psi0, Y0, Z0, fast=False):
        dev = self.device
        clk = self._clocks
        ev = clk["sample"].start()
        pt = device_perm(self.perm, dev)

        def perm_rows(X):
            X = fromnumpy(X, dev) if not isinstance(X, torch.Tensor) else X.to(dev, torch.float64)
            return X.contiguous() if pt is None else X[pt].contiguous()

        Om, Psi, Y, Z = (perm_rows(X) for X in (Om0, Psi0, Y0, Z0))
        self.nSamples = Om.shape[1]
        clk["sample"].stop(ev)
        clk["compress"].set(0.0)
        ev = clk["compress"].start()
        self._h = DeviceHBS.from_sketches(Om, Psi, Y, Z, self.Nb, self.L, rk, self.fac, pt)
        self.Nbvec = self._h.Nbvec
        clk["compress"].stop(ev)

    def constructHBS_ULV(self, rk, fast=False):
        _HBSnewMAT.constructHBS(self, rk)
        self.compute_ULV()

    def solve(self, b, device=None, mode="N"):
        return _HBSnewMAT.solve(self, b, mode)
