"""HBS matrices with ULV solves on the MI355X.

Mirror of matAssembly/HBS/HBSnew.py (SimonDirckx/SslabLU): class HBSMAT with the same
constructor, construct(rk, compute_ULV), matmat / rmatmat / matvec / rmatvec / @ / T,
solve(b, mode), set_mats, set_Nbvec, compute_ULV, nbytes and the attributes Umats,
Vmats, Dmats (stacked "reduced form" 2-D views), Nbvec, perm, fac, Nb, nl, L, shape,
dtype, nSamples, tSample, tCompress.  The numerics run in sslablu_amd.hbs.DeviceHBS
(batched MFMA GEMMs and QR kernels of libsslab_hip.so).

The operator A may be a float64 device tensor (sketched with MFMA GEMMs), an object
whose matmat/rmatmat take and return device tensors (attribute ``device_op = True``),
or any scipy-style LinearOperator / numpy array (sketched on the host, then uploaded).
Sketches use numpy.random.default_rng(seed) (seed attribute, default 0) where the
reference draws unseeded samples (HBSnew.py:218-219).
"""
from __future__ import annotations

import time

import numpy as np
import torch

from ... import _lib
from ...hbs import DeviceHBS, F64

fromnumpy = lambda X, dev: torch.as_tensor(np.ascontiguousarray(X, dtype=np.float64)).to(dev)


def _apply_op(A, X, transpose, device):
    """A X or A^T X for a device tensor X (N x s)."""
    if isinstance(A, torch.Tensor):
        return _dense_apply(A.to(device=device, dtype=F64), X, transpose)
    if getattr(A, "device_op", False):
        return A.rmatmat(X) if transpose else A.matmat(X)
    Xh = X.detach().cpu().numpy()
    Yh = A.rmatmat(Xh) if transpose else A.matmat(Xh)
    return fromnumpy(Yh, device)


def _dense_apply(Ad, X, transpose):
    from ...ops import dgemm
    return dgemm(Ad.t() if transpose else Ad, X)


class HBSMAT:
    """HBS approximation of a square operator on a cluster tree (HBSnew.py:65-416)."""

    def __init__(self, A=None, tree=None, quad=True, device=None, seed=0, sketch_fp32=False):
        """sketch_fp32: for a dense A, form the sketches Y = A Omega, Z = A^T Psi with the
        FP32 MFMA GEMM (A and the test matrices rounded to float); the compression that
        follows (QR, interpolative bases, null-space projections) stays FP64.  Linear
        operators are always applied in FP64."""
        _lib.require_gpu()
        self.sketch_fp32 = bool(sketch_fp32)
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.mode = "N"
        self._h = None
        self.Qlist, self.Rlist, self.Wlist, self.Uulist = [], [], [], []
        self.seed = seed
        if A is not None:
            self.A = A
            self.shape = tuple(A.shape)
            self.dtype = np.float64
        if tree is not None:
            self.perm = np.asarray(tree.perm_leaf)
            self.Nb = tree.nleaves
            if A is None:
                self.shape = (len(self.perm), len(self.perm))
            self.nl = self.shape[0] // self.Nb
            self.L = tree.nlevels
            self.tree = tree
        self.quad = quad
        self.fac = 4 if quad else 2
        self.blockSolveTime = self.nullTime = self.setupTime = self.DTime = 0.0
        self.tSample = 0.0
        self.tConstruct = 0.0
        self.tCompress = 0.0
        self.Nbvec = []

    # -- set-up --------------------------------------------------------------------
    def set_Nbvec(self, Nbvec):
        """Use a given block count per level with the identity permutation
        (HBSnew.py:119-124)."""
        self.Nbvec = list(Nbvec)
        self.Nb = Nbvec[0]
        self.nl = self.shape[0] // self.Nb
        self.L = len(Nbvec)
        self.perm = np.arange(self.shape[0], dtype=np.int64)

    def set_mats(self, Umats, Dmats, Vmats, Nbvec, fac=4):
        """Adopt given reduced-form factors (HBSnew.py:125-133): Umats[l] is
        (Nb_l*n_l, k), Dmats[l] (Nb_l*n_l, n_l)."""
        dev = self.device
        def blocks(X, nb):
            X = fromnumpy(X, dev) if not isinstance(X, torch.Tensor) else X.to(dev, F64)
            return X.reshape(nb, -1, X.shape[1]).contiguous()
        U = [blocks(u, Nbvec[i]) for i, u in enumerate(Umats)]
        V = [blocks(v, Nbvec[i]) for i, v in enumerate(Vmats)]
        D = [blocks(d, Nbvec[i]) for i, d in enumerate(Dmats)]
        self.fac = fac
        self._h = DeviceHBS(U, D, V, fac)
        self.Nbvec = list(Nbvec)
        self.Nb = Nbvec[0]
        self.perm = np.arange(self._h.N)
        self.shape = (self._h.N, self._h.N)
        self.dtype = np.float64

    @property
    def T(self):
        view = object.__new__(self.__class__)
        view.__dict__ = self.__dict__.copy()
        view.mode = "T" if self.mode == "N" else "N"
        return view

    # -- construction (HBSnew.py:140-295) ---------------------------------------------
    def construct(self, rk, compute_ULV=False):
        if compute_ULV:
            self.constructHBS_ULV(rk)
        else:
            self.constructHBS(rk)

    def _samples(self, rk):
        return 6 * rk + 4 * self.nl + 5 if self.fac == 4 else 4 * rk + 2 * self.nl + 5

    def constructHBS(self, rk):
        N = self.shape[0]
        s = self._samples(rk)
        self.nSamples = s
        dev = self.device
        t0 = time.time()
        rng = np.random.default_rng(self.seed)
        Om = rng.standard_normal((N, s))
        Psi = rng.standard_normal((N, s))
        perm = np.asarray(self.perm)
        Omp = np.zeros_like(Om)
        Omp[perm] = Om
        Psp = np.zeros_like(Psi)
        Psp[perm] = Psi
        pt = torch.as_tensor(perm).to(dev)
        if self.sketch_fp32 and isinstance(self.A, torch.Tensor):
            from ...ops import sgemm
            f32 = torch.float32
            A32 = self.A.to(device=dev, dtype=f32)
            Y = sgemm(A32, fromnumpy(Omp, dev).to(f32)).to(F64)[pt].contiguous()
            Z = sgemm(A32.t(), fromnumpy(Psp, dev).to(f32)).to(F64)[pt].contiguous()
            del A32
        else:
            Y = _apply_op(self.A, fromnumpy(Omp, dev), False, dev)[pt].contiguous()
            Z = _apply_op(self.A, fromnumpy(Psp, dev), True, dev)[pt].contiguous()
        torch.cuda.synchronize(dev)
        self.tSample += time.time() - t0
        t0 = time.time()
        self._h = DeviceHBS.from_sketches(fromnumpy(Om, dev), fromnumpy(Psi, dev), Y, Z,
                                          self.Nb, self.L, rk, self.fac, perm)
        self.Nbvec = self._h.Nbvec
        torch.cuda.synchronize(dev)
        self.tCompress = time.time() - t0

    def constructHBS_ULV(self, rk):
        self.constructHBS(rk)
        self.compute_ULV()

    def compute_ULV(self):
        t0 = time.time()
        lev = self._h.ulv_factor("N")
        self.Qlist = [l["Q"].reshape(-1, l["Q"].shape[2]) if not l["root"] else None for l in lev]
        self.Wlist = [l["W"].reshape(-1, l["W"].shape[2]) for l in lev if not l["root"]]
        self.Uulist = [l["Uu"] for l in lev if not l["root"]]
        self.Rlist = [l["R11inv"] for l in lev if not l["root"]]
        self.tCompress += time.time() - t0

    # -- attributes in the reference's stacked 2-D form --------------------------------
    @property
    def Umats(self):
        return [] if self._h is None else [U.reshape(-1, U.shape[2]) for U in self._h.Umats]

    @property
    def Vmats(self):
        return [] if self._h is None else [V.reshape(-1, V.shape[2]) for V in self._h.Vmats]

    @property
    def Dmats(self):
        return [] if self._h is None else [D.reshape(-1, D.shape[2]) for D in self._h.Dmats]

    @property
    def nbytes(self):
        return self._h.nbytes if self._h is not None else 0

    # -- apply (HBSnew.py:303-388) -----------------------------------------------------
    def _wrap(self, v, fn):
        was_np = not isinstance(v, torch.Tensor)
        x = fromnumpy(v, self.device) if was_np else v
        y = fn(x)
        return y.cpu().numpy() if was_np else y

    def matmat(self, v):
        return self._wrap(v, lambda x: self._h.apply(x, "N"))

    def rmatmat(self, v):
        return self._wrap(v, lambda x: self._h.apply(x, "T"))

    def matvec(self, v):
        return self.matmat(v)

    def rmatvec(self, v):
        return self.rmatmat(v)

    def __matmul__(self, v):
        if self.mode == "N":
            return self.matmat(v)
        if self.mode == "T":
            return self.rmatmat(v)
        raise ValueError("mode not recognized")

    # -- solve (HBSnew.py:393-416) -----------------------------------------------------
    def solve(self, b, mode="N"):
        if mode not in ("N", "T"):
            raise NotImplementedError("mode not recognized")
        if self.mode == "T":
            mode = "T" if mode == "N" else "N"
        return self._wrap(b, lambda x: self._h.solve(x, mode))
