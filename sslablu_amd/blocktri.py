"""Device-resident block-tridiagonal S-system factorization (FP64, gfx950).

    T_i x_i + SiM_i x_{i-1} + SiP_i x_{i+1} = b_i,   i = 0..nb-1

`BlockTriFactor` owns the factor tensors (torch, on the GPU) and calls
libsslab_hip.so for the factorization and the multi-RHS solves:

* method "rb": red-black cyclic reduction (any nb; log2(nb) levels, every level one
  batched MFMA GEMM launch per sweep direction);
* method "thomas": sequential block elimination (direct_solve/omsdirectsolve.py:48-141).

Periodic (cyclic) systems add the corner blocks E (row 0, column nb-1) and F (row
nb-1, column 0) through a rank-2m Woodbury correction as in
direct_solve/omsdirectsolve.py:143-175, but stored as Z = T^{-1}U (N x 2m) and
K = (I + V^T Z)^{-1} (2m x 2m) instead of the dense N x N `smw_block`:
x = y - Z K [y_{nb-1}; y_0].
"""
from __future__ import annotations

import ctypes
import warnings

import torch

from . import _lib, tuning
from ._lib import call, ptr, stream_handle

F64 = torch.float64


def _dev_blocks(x, nb, m, device):
    """Stack a list of m x m blocks (numpy / torch) into a contiguous (nb, m, m)
    float64 device tensor."""
    if isinstance(x, torch.Tensor) and x.dim() == 3:
        t = x.to(device=device, dtype=F64)
    else:
        t = torch.stack([torch.as_tensor(b, dtype=F64) for b in x]).to(device)
    t = t.contiguous()
    if tuple(t.shape) != (nb, m, m):
        raise ValueError(f"expected {nb} blocks of {m}x{m}, got {tuple(t.shape)}")
    return t


class BlockTriFactor:
    """Factorization of one block-tridiagonal system on the GPU."""

    def __init__(self, method: str = "rb", device=None):
        if method not in ("rb", "thomas"):
            raise ValueError(f"unknown method {method!r}")
        _lib.require_gpu()
        _lib.load()
        self.method = method
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.nb = self.m = 0
        self.factor = None
        self.info = None
        self.cyclic = False
        self.ident0 = False  # red-black factor of identity diagonal blocks (T = None)
        self.chained_backward = False  # backward sweep as one dependency-ordered launch
        self.Z = None
        self.K = None

    # -- factorization ---------------------------------------------------------
    def factorize(self, T, SiM, SiP, E=None, F=None, stream=None):
        """T, SiM, SiP: (nb, m, m) blocks (SiM[0] and SiP[nb-1] are ignored); T = None
        means identity diagonal blocks (the oms S-system; red-black then skips the
        first level's LU).  E, F: optional corner blocks for a periodic system."""
        ref = T if T is not None else SiM
        nb = len(ref)
        m = ref[0].shape[0]
        dev = self.device
        self.ident0 = T is None and self.method == "rb"
        with torch.cuda.device(dev):
            if T is None and self.method == "thomas":
                T = torch.eye(m, dtype=F64, device=dev).expand(nb, m, m)
            Td = _dev_blocks(T, nb, m, dev) if T is not None else None
            Md = _dev_blocks(SiM, nb, m, dev)
            Pd = _dev_blocks(SiP, nb, m, dev)
            nfac, nwork, niwork = _lib.sizes(self.method, nb, m)
            self.factor = torch.empty(nfac, dtype=F64, device=dev)
            work = torch.empty(max(nwork, 1), dtype=F64, device=dev)
            iwork = torch.empty(max(niwork, 1), dtype=torch.int32, device=dev)
            ninfo = _lib.load().sslab_rb_info_count(nb) if self.method == "rb" else nb
            info = torch.zeros(ninfo, dtype=torch.int32, device=dev)
            s = stream_handle(stream)
            call(f"sslab_{self.method}_factor", nb, m, ptr(Td) if Td is not None else None,
                 ptr(Md), ptr(Pd), ptr(self.factor), ptr(work), ptr(iwork), ptr(info), s)
            self.nb, self.m = nb, m
            if self.method == "thomas":
                # keep the pivots of LU(B_i) for the reference's `B` attribute
                self.ipiv = iwork[: nb * m].view(nb, m).clone()
            self.info = info
            bad = int((info != 0).sum().item())
            if bad:
                warnings.warn(f"{bad} exactly singular pivot block(s) in the block-tridiagonal "
                              "factorization; results will contain inf/nan", RuntimeWarning)
            self.cyclic = E is not None or F is not None
            if self.cyclic:
                self._factor_woodbury(E, F, stream)
        return self

    def _factor_woodbury(self, E, F, stream):
        nb, m, dev = self.nb, self.m, self.device
        if nb < 2:
            raise ValueError("a cyclic block system needs at least two blocks")
        Ed = torch.as_tensor(E, dtype=F64).to(dev) if E is not None else torch.zeros(m, m, dtype=F64, device=dev)
        Fd = torch.as_tensor(F, dtype=F64).to(dev) if F is not None else torch.zeros(m, m, dtype=F64, device=dev)
        U = torch.zeros(nb * m, 2 * m, dtype=F64, device=dev)
        U[0:m, 0:m] = Ed
        U[(nb - 1) * m:, m:] = Fd
        self.cyclic = False  # plain solve for Z
        Z = self.solve(U, stream=stream)
        self.cyclic = True
        # G = I + V^T Z with V^T Z = [Z_{nb-1}; Z_0]
        G = torch.eye(2 * m, dtype=F64, device=dev)
        G[0:m] += Z[(nb - 1) * m:]
        G[m:] += Z[0:m]
        K = torch.empty_like(G)
        ipiv = torch.empty(2 * m, dtype=torch.int32, device=dev)
        perm = torch.empty(2 * m, dtype=torch.int32, device=dev)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        eye = torch.eye(2 * m, dtype=F64, device=dev)
        s = stream_handle(stream)
        call("sslab_dgetrf_batched", 2 * m, ptr(G), 2 * m, 0, ptr(ipiv), 0, ptr(info), 1, s)
        call("sslab_dgetrs_batched", 2 * m, 2 * m, ptr(G), 2 * m, 0, ptr(ipiv), 0, ptr(eye),
             2 * m, 0, ptr(K), 2 * m, 0, ptr(perm), 1, s)
        self.Z, self.K = Z, K

    # -- solves -----------------------------------------------------------------
    def solve_(self, V: torch.Tensor | None, X: torch.Tensor, stream=None,
               level0_reduced: bool = False, phase: str = "all"):
        """Solve in place of preallocated device buffers: V (N, r) holds the
        right-hand sides and is overwritten, X (N, r) receives the solution.
        Both float64, row-major with unit column stride.  level0_reduced: V and X
        hold the entry state written by HPSSlabOperator.rhs_rb0_from_boundary (V's even
        block rows reduced by the first red-black level, X's odd block rows the odd
        right-hand sides, X's even rows zero); red-black factors built with T = None
        only.  phase "forward": the forward sweep and coarsest level only, leaving the
        backward sweep's entry state in X; "backward": X holds that state (V unused, may be
        None) -- red-black factors only.  With the attribute chained_backward set, a
        red-black backward sweep runs as one dependency-ordered launch
        (SSLAB_RB_CHAINED_BACKWARD); chained solves on one GPU must not overlap in time."""
        nb, m = self.nb, self.m
        N = nb * m
        if phase not in ("all", "forward", "backward"):
            raise ValueError(f"solve_: unknown phase {phase!r}")
        if phase != "all" and (self.method != "rb" or self.cyclic):
            raise ValueError("solve_: split phases need a non-cyclic red-black factor")
        if phase == "backward":
            if level0_reduced:
                raise ValueError("solve_: backward phase excludes level0_reduced")
            V = X  # shape / dtype checks only: the library does not read it
        if self.factor is None:
            raise ValueError("solve_: no FP64 factor (factorize_mixed keeps only the FP32 one; "
                             "use solve_mixed_)")
        if V.shape[0] != N or X.shape[0] != N or V.dim() != 2 or X.shape[1] != V.shape[1]:
            raise ValueError("solve_: V and X must be (nb*m, r)")
        if V.dtype != F64 or X.dtype != F64 or (V.shape[1] > 0 and (V.stride(1) != 1 or
                                                                    X.stride(1) != 1)):
            raise ValueError("solve_: float64 buffers with unit column stride required")
        if level0_reduced and not (self.method == "rb" and self.ident0 and not self.cyclic
                                   and nb >= 2):
            raise ValueError("solve_: level0_reduced needs a red-black factor with T = None "
                             "and at least two blocks")
        r = V.shape[1]
        if r == 0:
            return X
        s = stream_handle(stream)
        flags = ((1 if self.ident0 else 0) | (2 if level0_reduced else 0) |
                 {"all": 0, "forward": 4, "backward": 8}[phase] |
                 (16 if getattr(self, "chained_backward", False) else 0))
        bwd = getattr(self, "_bwd_policies", {}).get(r) if self.method == "rb" else None
        with tuning.policy(tuning.lookup(self.tune_key(r))), tuning.backward_policies(bwd):
            if self.method == "rb":
                call("sslab_rb_solve_ex", nb, m, r, ptr(self.factor),
                     0 if phase == "backward" else ptr(V), V.stride(0),
                     ptr(X), X.stride(0), flags, s)
            else:
                call("sslab_thomas_solve", nb, m, r, ptr(self.factor), ptr(V), V.stride(0),
                     ptr(X), X.stride(0), s)
        if self.cyclic:
            # t = K [x_{nb-1}; x_0];  x -= Z t
            t = torch.empty(2 * m, r, dtype=F64, device=X.device)
            K = self.K
            call("sslab_dgemm_strided_batched", 0, 0, 2 * m, r, m, 1.0, ptr(K), 2 * m, 0,
                 ptr(X[(nb - 1) * m:]), X.stride(0), 0, 0.0, ptr(t), r, 0, 1, s)
            call("sslab_dgemm_strided_batched", 0, 0, 2 * m, r, m, 1.0, ptr(K[:, m:]), 2 * m, 0,
                 ptr(X), X.stride(0), 0, 1.0, ptr(t), r, 0, 1, s)
            call("sslab_dgemm_strided_batched", 0, 0, N, r, 2 * m, -1.0, ptr(self.Z), 2 * m, 0,
                 ptr(t), r, 0, 1.0, ptr(X), X.stride(0), 0, 1, s)
        return X

    def solve(self, B: torch.Tensor, stream=None) -> torch.Tensor:
        """Return X with X = S^{-1} B for a device tensor B of shape (N,) or (N, r)."""
        one_d = B.dim() == 1
        V = B.reshape(B.shape[0], -1).to(device=self.device, dtype=F64).clone(memory_format=torch.contiguous_format)
        X = torch.empty_like(V)
        self.solve_(V, X, stream)
        return X[:, 0] if one_d else X

    # -- mixed precision --------------------------------------------------------------
    def enable_mixed(self, T=None, SiM=None, SiP=None, stream=None):
        """Keep an FP32 copy of the red-black factor (the refinement preconditioner) and
        the FP64 blocks the refinement residual needs (T = None: identity diagonal)."""
        if self.method != "rb":
            raise ValueError("mixed precision uses the red-black factorization")
        if self.cyclic:
            raise NotImplementedError("mixed-precision solves of periodic systems")
        self.factor32 = torch.empty(self.factor.numel(), dtype=torch.float32, device=self.device)
        call("sslab_cast_f64_to_f32", self.factor.numel(), ptr(self.factor), ptr(self.factor32),
             stream_handle(stream))
        nb, m = self.nb, self.m
        self._T64 = None if T is None else _dev_blocks(T, nb, m, self.device)
        self._M64 = _dev_blocks(SiM, nb, m, self.device)
        self._P64 = _dev_blocks(SiP, nb, m, self.device)
        return self

    def factorize_mixed(self, T, SiM, SiP, stream=None):
        """Factor the system in FP32 (FP32 LU and FP32 MFMA Schur updates of the red-black
        elimination) as the low-precision preconditioner of solve_mixed_; the FP64 blocks
        are kept for the refinement residual.  T = None: identity diagonal blocks.  Only
        solve_mixed_ is available afterwards (there is no FP64 factor)."""
        if self.method != "rb":
            raise ValueError("mixed precision uses the red-black factorization")
        ref = T if T is not None else SiM
        nb = len(ref)
        m = ref[0].shape[0]
        dev = self.device
        with torch.cuda.device(dev):
            Td = _dev_blocks(T, nb, m, dev) if T is not None else None
            Md = _dev_blocks(SiM, nb, m, dev)
            Pd = _dev_blocks(SiP, nb, m, dev)
            s = stream_handle(stream)

            def f32(x):
                y = torch.empty(x.shape, dtype=torch.float32, device=dev)
                call("sslab_cast_f64_to_f32", x.numel(), ptr(x), ptr(y), s)
                return y

            T32 = f32(Td) if Td is not None else None
            M32, P32 = f32(Md), f32(Pd)
            nfac, nwork, niwork = _lib.sizes("rb", nb, m)
            self.factor32 = torch.empty(nfac, dtype=torch.float32, device=dev)
            work = torch.empty(max(nwork, 1), dtype=torch.float32, device=dev)
            iwork = torch.empty(max(niwork, 1), dtype=torch.int32, device=dev)
            info = torch.zeros(_lib.load().sslab_rb_info_count(nb), dtype=torch.int32, device=dev)
            call("sslab_rb_factor_f32", nb, m, ptr(T32) if T32 is not None else None, ptr(M32),
                 ptr(P32), ptr(self.factor32), ptr(work), ptr(iwork), ptr(info), s)
            self.nb, self.m = nb, m
            self.factor = None
            self.info = info
            self.cyclic = False
            self._T64, self._M64, self._P64 = Td, Md, Pd
            bad = int((info != 0).sum().item())
            if bad:
                warnings.warn(f"{bad} exactly singular pivot block(s) in the FP32 block-tridiagonal "
                              "factorization", RuntimeWarning)
        return self

    def solve_mixed_(self, B: torch.Tensor, X: torch.Tensor, iters: int = 2, stream=None,
                     work=None):
        """X = S^{-1} B by FP32 red-black solves refined with FP64 residuals
        (`iters` refinement steps).  B, X: contiguous float64 (N, r) device tensors."""
        nb, m = self.nb, self.m
        N, r = B.shape
        if N != nb * m or tuple(X.shape) != (N, r) or not (B.is_contiguous() and X.is_contiguous()):
            raise ValueError("solve_mixed_: contiguous (nb*m, r) float64 B and X required")
        if work is None:
            work = (torch.empty(N * r, dtype=F64, device=self.device),
                    torch.empty(2 * N * r, dtype=torch.float32, device=self.device))
        w64, w32 = work
        with tuning.policy(tuning.lookup(self.tune_key(r, "mixed"))):
            call("sslab_rb_solve_mixed", nb, m, r, ptr(self.factor32),
                 ptr(self._T64) if self._T64 is not None else None, ptr(self._M64),
                 ptr(self._P64), ptr(B), ptr(X), ptr(w64), ptr(w32), int(iters),
                 stream_handle(stream))
        return X

    def solve_mixed_noalloc_(self, B: torch.Tensor, X: torch.Tensor, iters: int = 2,
                             stream=None):
        """solve_mixed_ without a workspace (sslab_rb_solve_mixed_noalloc): the library's
        static per-GPU scratch holds the FP64 residual and the FP32 panels, so the call
        allocates nothing (usable under HIP graph capture); 16 * nb*m*r bytes must fit its
        64 MiB.  Calls that overlap on one GPU race on that scratch."""
        nb, m = self.nb, self.m
        N, r = B.shape
        if N != nb * m or tuple(X.shape) != (N, r) or not (B.is_contiguous() and X.is_contiguous()):
            raise ValueError("solve_mixed_noalloc_: contiguous (nb*m, r) float64 B and X required")
        with tuning.policy(tuning.lookup(self.tune_key(r, "mixed"))):
            call("sslab_rb_solve_mixed_noalloc", nb, m, r, ptr(self.factor32),
                 ptr(self._T64) if self._T64 is not None else None, ptr(self._M64),
                 ptr(self._P64), ptr(B), ptr(X), int(iters), stream_handle(stream))
        return X

    # -- autotuning ---------------------------------------------------------------------
    def tune_key(self, r: int, precision: str = "fp64") -> str:
        return f"blocktri:{self.method}:{self.nb}:{self.m}:{r}:{precision}"

    def autotune(self, r: int, precision: str = "fp64", iters: int = 1, reps: int = 10,
                 fn=None, cands=None):
        """Pick the GEMM tile policy for solves with `r` right-hand sides by timing every
        candidate on this factor; later solve_ / solve_mixed_ calls of that width use
        it.  `fn` (optional) is the caller's own step to time instead of a bare solve
        (e.g. right-hand-side assembly plus solve_); it runs under each candidate policy.
        Returns (policy, microseconds per call)."""
        N = self.nb * self.m
        V = torch.randn(N, r, dtype=F64, device=self.device)
        X = torch.empty_like(V)
        if fn is not None:
            pass
        elif precision == "mixed":
            work = (torch.empty(N * r, dtype=F64, device=self.device),
                    torch.empty(2 * N * r, dtype=torch.float32, device=self.device))
            fn = lambda: self.solve_mixed_(V, X, iters=iters, work=work)  # noqa: E731
        else:
            V0 = V.clone()
            fn = lambda: (V.copy_(V0), self.solve_(V, X))  # noqa: E731
        key = self.tune_key(r, precision)
        with tuning.policy(tuning.DEFAULT):
            tuning.lookup(key)  # load a persisted cache first, then drop a stale choice
            _saved = tuning._cache.pop(key, None)
            try:
                return tuning.autotune(fn, key, cands=cands, reps=reps)
            except Exception:
                if _saved is not None:
                    tuning._cache[key] = _saved
                raise

    def autotune_backward_steps(self, r: int, reps: int = 10, cands=None, fn=None):
        """Pick a tile policy per step of the red-black backward sweep (its levels differ in
        how many blocks they multiply) by timing each step's candidates with the other steps
        fixed, starting from the policy tuned for the whole solve; later solve_ calls of
        width r use the table.  `fn` (optional) is the caller's step to time (default: the
        backward sweep alone).  Returns (list of policies, microseconds per call)."""
        if self.method != "rb":
            raise ValueError("autotune_backward_steps: red-black factors only")
        _lib.require_gpu()
        levels, _ = rb_levels(self.nb, self.m)
        if levels == 0:
            return [], 0.0
        X = torch.randn(self.nb * self.m, r, dtype=F64, device=self.device)
        if fn is None:
            fn = lambda: self.solve_(None, X, phase="backward")  # noqa: E731
        cands = list(cands) if cands is not None else tuning.candidates()
        base = tuning.lookup(self.tune_key(r)) or tuning.get_policy()
        table = [base] * levels
        store = self.__dict__.setdefault("_bwd_policies", {})
        store.pop(r, None)
        s = torch.cuda.current_stream()

        def timed():
            fn()
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(reps):
                fn()
            e1.record(s)
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) * 1e3 / reps

        best_t = None
        for u in range(levels):
            for p in [table[u]] + cands:
                trial = table[:u] + [p] + table[u + 1:]
                with tuning.backward_policies(trial):
                    t = timed()
                if best_t is None or t < best_t:
                    best_t, table = t, trial
        store[r] = table
        return table, best_t

    def residual(self, B: torch.Tensor, X: torch.Tensor, stream=None) -> torch.Tensor:
        """R = B - S X in FP64 (requires enable_mixed for the stored blocks)."""
        R = torch.empty_like(B)
        nb, m = self.nb, self.m
        r = B.shape[1]
        call("sslab_rb_residual", nb, m, r, ptr(self._T64) if self._T64 is not None else None,
             ptr(self._M64), ptr(self._P64), ptr(B), r, ptr(X), r, ptr(R), r,
             stream_handle(stream))
        return R

    # -- introspection ----------------------------------------------------------
    @property
    def nbytes(self) -> int:
        n = self.factor.numel() * 8 if self.factor is not None else 0
        if getattr(self, "factor32", None) is not None:
            n += self.factor32.numel() * 4
        if self.Z is not None:
            n += (self.Z.numel() + self.K.numel()) * 8
        return n


def rb_levels(nb: int, m: int) -> tuple[int, int]:
    """(elimination levels, nodes of the coarsest level) of the red-black factor of nb
    blocks of size m: at least one level when nb >= 2, then halving while more than
    min(4, 2048 // m) nodes remain; the coarsest level is stored as a dense inverse."""
    top_max = max(1, min(4, 2048 // m))
    n, L = nb, 0
    while n > 1 and (L == 0 or n > top_max):
        n, L = (n + 1) // 2, L + 1
    return L, n


def rb_backward_blocks(nb: int, m: int) -> int:
    """m x m blocks one right-hand-side column of the red-black backward sweep multiplies
    (x_j -= (T^-1 SiM)_j x_{j-1} + (T^-1 SiP)_j x_{j+1} over every level's odd nodes)."""
    L, _ = rb_levels(nb, m)
    n, blocks = nb, 0
    for _ in range(L):
        nO = n // 2
        nboth = nO - 1 if n % 2 == 0 else nO
        blocks += 2 * nboth + (nO - nboth)
        n = (n + 1) // 2
    return blocks


def solve_multi(factors, Vs, Xs, stream=None):
    """Solve several red-black-factored systems together (sslab_rb_solve_multi): every
    elimination level of the systems with at most 32 right-hand sides shares one grouped
    GEMM launch, so many small systems of different block sizes go through the sweeps at
    once; wider systems keep their own launches, which fill the chip by themselves.
    Vs[i] (overwritten) / Xs[i]: float64 device tensors (nb_i*m_i, r_i), unit column
    stride, as for BlockTriFactor.solve_."""
    import ctypes
    n = len(factors)
    if not (len(Vs) == len(Xs) == n):
        raise ValueError("solve_multi: one V and one X per factor")
    for f, V, X in zip(factors, Vs, Xs):
        if f.method != "rb" or f.cyclic:
            raise ValueError("solve_multi: non-cyclic red-black factors only")
        N = f.nb * f.m
        if V.shape[0] != N or X.shape[0] != N or V.dim() != 2 or X.shape[1] != V.shape[1]:
            raise ValueError("solve_multi: V and X must be (nb*m, r)")
        if V.dtype != F64 or X.dtype != F64 or (V.shape[1] > 0 and (V.stride(1) != 1 or
                                                                    X.stride(1) != 1)):
            raise ValueError("solve_multi: float64 buffers with unit column stride required")
    ints = lambda xs: (ctypes.c_int * n)(*xs)  # noqa: E731
    ptrs = lambda ts: (ctypes.c_void_p * n)(*[ptr(t) for t in ts])  # noqa: E731
    nb = ints([f.nb for f in factors])
    m = ints([f.m for f in factors])
    r = ints([V.shape[1] for V in Vs])
    fp = ptrs([f.factor for f in factors])
    vp = ptrs(Vs)
    xp = ptrs(Xs)
    ldv = ints([max(V.stride(0), V.shape[1]) for V in Vs])
    ldx = ints([max(X.stride(0), X.shape[1]) for X in Xs])
    A = ctypes.addressof
    call("sslab_rb_solve_multi", n, A(nb), A(m), A(r), A(fp), A(vp), A(ldv), A(xp), A(ldx),
         stream_handle(stream))
    return Xs


def chain_status() -> bool:
    """True when a chained backward sweep on the current GPU gave up waiting since the
    last call (its solve is then invalid); synchronises the device and clears the flag
    (sslab_rb_chain_status)."""
    v = ctypes.c_int(0)
    call("sslab_rb_chain_status", ctypes.addressof(v))
    return bool(v.value)
