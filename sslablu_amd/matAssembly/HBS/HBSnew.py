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
Gaussian test matrices are drawn on the device from a torch.Generator seeded with the
seed attribute (default 0) where the reference draws unseeded numpy samples
(HBSnew.py:218-219); nothing crosses PCIe for a device operator.
"""
from __future__ import annotations


import numpy as np
import torch

from ... import _lib
from ...hbs import DeviceHBS, F64, device_perm

fromnumpy = lambda X, dev: torch.as_tensor(np.ascontiguousarray(X, dtype=np.float64)).to(dev)


class _StageClock:
    """Seconds spent in a construction stage, measured with HIP events on the current
    stream: no host synchronisation while the stages run (the slab eliminations queue
    the next compression behind this one); the events are read, and waited for, only
    when the time is asked for."""

    def __init__(self):
        self.base = 0.0
        self.open = []

    def start(self):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def stop(self, e0):
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        self.open.append((e0, e1))

    def value(self):
        for e0, e1 in self.open:
            e1.synchronize()
            self.base += e0.elapsed_time(e1) / 1e3
        self.open = []
        return self.base

    def set(self, v):
        self.base = float(v)
        self.open = []


def _stage_time(name):
    def get(self):
        return self._clocks[name].value()

    def put(self, v):
        self._clocks[name].set(v)
    return property(get, put)


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

    tSample = _stage_time("sample")
    tCompress = _stage_time("compress")

    def __init__(self, A=None, tree=None, quad=True, device=None, seed=0, sketch_fp32=False,
                 samples=None):
        """sketch_fp32: for a dense A, form the sketches Y = A Omega, Z = A^T Psi with the
        FP32 MFMA GEMM (A and the test matrices rounded to float); the compression that
        follows (QR, interpolative bases, null-space projections) stays FP64.  Linear
        operators are always applied in FP64.
        samples: sketch columns per side, an int or a callable rk -> int; None keeps the
        reference's rule (HBSnew.py:147-150: 4 rk + 2 leaf + 5 binary, 6 rk + 4 leaf + 5
        quad)."""
        _lib.require_gpu()
        self.sketch_fp32 = bool(sketch_fp32)
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.mode = "N"
        self._h = None
        self.Qlist, self.Rlist, self.Wlist, self.Uulist = [], [], [], []
        self.seed = seed
        if samples is not None and not callable(samples) and int(samples) < 1:
            raise ValueError("samples must be a positive int or a callable rk -> int")
        self.samples = samples
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
        self._clocks = {"sample": _StageClock(), "compress": _StageClock()}
        self.tConstruct = 0.0
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
        if self.samples is not None:
            return int(self.samples(rk)) if callable(self.samples) else int(self.samples)
        return 6 * rk + 4 * self.nl + 5 if self.fac == 4 else 4 * rk + 2 * self.nl + 5

    def constructHBS(self, rk):
        Om, Psi, Y, Z, pt = self._sketch(rk)
        clk = self._clocks
        clk["compress"].set(0.0)
        ev = clk["compress"].start()
        self._h = DeviceHBS.from_sketches(Om, Psi, Y, Z, self.Nb, self.L, rk, self.fac, pt)
        self.Nbvec = self._h.Nbvec
        clk["compress"].stop(ev)

    def _sketch(self, rk):
        """Seeded test matrices and the sketches Y = A Om, Z = A^T Psi in leaf order
        (HBSnew.py:147-150, 218-225); returns (Om, Psi, Y, Z, device permutation)."""
        N = self.shape[0]
        s = self._samples(rk)
        self.nSamples = s
        dev = self.device
        clk = self._clocks
        ev = clk["sample"].start()
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(self.seed))
        Om = torch.randn(N, s, generator=gen, device=dev, dtype=F64)
        Psi = torch.randn(N, s, generator=gen, device=dev, dtype=F64)
        pt = device_perm(self.perm, dev)      # None: identity leaf order, no gathers
        if pt is None:
            Omp, Psp = Om, Psi
            rows = lambda X: X.contiguous()
        else:
            Omp = torch.empty_like(Om)
            Omp[pt] = Om
            Psp = torch.empty_like(Psi)
            Psp[pt] = Psi
            rows = lambda X: X[pt].contiguous()
        if self.sketch_fp32 and isinstance(self.A, torch.Tensor):
            from ...ops import sgemm
            f32 = torch.float32
            A32 = self.A.to(device=dev, dtype=f32)
            Y = rows(sgemm(A32, Omp.to(f32)).to(F64))
            Z = rows(sgemm(A32.t(), Psp.to(f32)).to(F64))
            del A32
        else:
            Y = rows(_apply_op(self.A, Omp, False, dev))
            Z = rows(_apply_op(self.A, Psp, True, dev))
        del Omp, Psp
        clk["sample"].stop(ev)
        return Om, Psi, Y, Z, pt

    def constructHBS_ULV(self, rk):
        self.constructHBS(rk)
        self.compute_ULV()

    def compute_ULV(self, both=False):
        """ULV factors for solve(mode="N"); with both=True those of the transpose too
        (mode "T", otherwise factored on its first use), from one stacked factorization."""
        ev = self._clocks["compress"].start()
        lev = self._h.ulv_factor_pair() if both else self._h.ulv_factor("N")
        self.Qlist = [l["Q"].reshape(-1, l["Q"].shape[2]) if not l["root"] else None for l in lev]
        self.Wlist = [l["W"].reshape(-1, l["W"].shape[2]) for l in lev if not l["root"]]
        self.Uulist = [l["Uu"] for l in lev if not l["root"]]
        self.Rlist = [l["R11inv"] for l in lev if not l["root"]]
        self._clocks["compress"].stop(ev)

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


def construct_many(mats, rk, compute_ULV=False):
    """Compress several HBSMATs over the same tree (same leaf order, sample count and
    rank) together: each operator is sketched on its own, then the sketches are stacked
    along every tree level's batch, so the compression's batched QRs and GEMMs (and the
    ULV factorization, compute_ULV: both modes) run once for all of them -- a batched
    launch costs about what one operator's costs at the upper tree levels, where each
    operator has only a few boxes.  Equivalent to construct(rk, compute_ULV) on each."""
    mats = list(mats)
    if not mats:
        return mats
    h0 = mats[0]
    sk = [h._sketch(rk) for h in mats]
    K = len(mats)
    for h in mats[1:]:
        if (h.Nb, h.L, h.fac, h.shape) != (h0.Nb, h0.L, h0.fac, h0.shape):
            raise ValueError("construct_many: every matrix needs the same tree and shape")
    ev = h0._clocks["compress"].start()
    big = DeviceHBS.from_sketches(*(torch.cat([s_[j] for s_ in sk]) for j in range(4)),
                                  h0.Nb * K, h0.L, rk, h0.fac, None)
    if compute_ULV:
        big.ulv_factor_pair()
    parts = big.split(K, perm=sk[0][4])
    for h, part in zip(mats, parts):
        h._h = part
        h.Nbvec = part.Nbvec
        h._clocks["compress"].set(0.0)
        if compute_ULV:
            h.compute_ULV()
    h0._clocks["compress"].stop(ev)
    return mats
