"""Device HBS matrices: randomized compression, telescoping apply and ULV solve.

An HBS (hierarchically block-separable, weak admissibility) matrix in "reduced form"
(matAssembly/HBS/HBSnew.py:65-139):

    A = D_0 + U_0 (D_1 + U_1 ( ... D_root ... ) V_1^T) V_0^T

with block-diagonal D_l (Nb_l blocks n_l x n_l), U_l, V_l (Nb_l blocks n_l x k), and
n_{l+1} = fac * k.  Every level is one strided batch of small dense blocks, so each step
of the compression (HBSnew.py:147-210), the apply (:303-388) and the ULV factorization
and solve (ULVsparse.py:231-358) is a handful of batched launches of libsslab_hip.so:
MFMA GEMMs, Householder / column-pivoted QR, triangular solves.

Differences from the reference, on purpose:
* the column basis of each box is taken from a column-pivoted QR of Y_i P_i (the
  rank-revealing FP64 pivot selection) instead of an SVD; for an exactly rank-k block
  both give the same subspace;
* the triangular blocks of the ULV factors are inverted once at factorization time so
  the solve is GEMMs only;
* transposed solves factor A^T = D^T + V (...)^T U^T on first use.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_handle

F64 = torch.float64
I32 = torch.int32

_PERMS: dict = {}
_PLANS = os.environ.get("SSLAB_HBS_PLANS", "1") != "0"


class _Recorder(threading.local):
    """Per-thread record of the GEMM launches of a DeviceHBS solve / apply being planned:
    (entry point, raw arguments, objects the arguments point into).  `bad` is set when a
    launch needed a copy of an operand (GPU work that a replay would not redo)."""
    ops = None
    bad = False


_rec = _Recorder()


def _recording_copy():
    if _rec.ops is not None:
        _rec.bad = True


def device_perm(perm, device):
    """Leaf permutation as an int64 device tensor, or None for the identity (no gathers).
    Host arrays are uploaded once per distinct permutation and device: every compression
    of a slab elimination reuses its tree's copy instead of a blocking pageable upload."""
    if perm is None:
        return None
    if isinstance(perm, torch.Tensor) and perm.is_cuda:
        return perm.to(device=device, dtype=torch.int64)
    p = np.ascontiguousarray(np.asarray(perm), dtype=np.int64)
    key = (str(torch.device(device)), p.tobytes())
    if key not in _PERMS:
        if len(_PERMS) >= 16:
            _PERMS.clear()
        ident = p.ndim == 1 and np.array_equal(p, np.arange(p.size))
        _PERMS[key] = None if ident else torch.as_tensor(p).to(device)
    return _PERMS[key]


# ---------------------------------------------------------------------------
# batched primitives on (nb, rows, cols) device tensors with unit last stride

def _bstride(X):
    return 0 if X.shape[0] == 1 else X.stride(0)


def _check_f64(where, *tensors):
    """Every operand handed to the FP64 kernels must be a float64 GPU tensor: the
    kernels take raw pointers and would read past a float32 or host allocation."""
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{where}: expected a torch.Tensor, got {type(t).__name__}")
        if t.dtype != F64:
            raise TypeError(f"{where}: float64 operands required, got {t.dtype}")
        if t.device.type != "cuda":
            raise TypeError(f"{where}: GPU operands required, got a tensor on {t.device}")


def bgemm(A, B, out=None, alpha=1.0, beta=0.0, ta=False, tb=False):
    """out_b = alpha op(A_b) op(B_b) + beta out_b for (nb, ., .) tensors (a batch of 1
    broadcasts)."""
    _check_f64("bgemm", A, B, out)
    if A.stride(2) != 1:
        _recording_copy()
        A = A.contiguous()
    if B.stride(2) != 1:
        _recording_copy()
        B = B.contiguous()
    nb = max(A.shape[0], B.shape[0])
    M = A.shape[2] if ta else A.shape[1]
    K = A.shape[1] if ta else A.shape[2]
    K2 = B.shape[2] if tb else B.shape[1]
    N = B.shape[1] if tb else B.shape[2]
    if K != K2:
        raise ValueError(f"bgemm: inner dimensions {K} != {K2}")
    if out is None:
        out = torch.empty(nb, M, N, dtype=F64, device=A.device)
        beta = 0.0
    if out.stride(2) != 1 or tuple(out.shape) != (nb, M, N):
        raise ValueError("bgemm: bad output")
    args = (int(ta), int(tb), M, N, K, float(alpha), ptr(A), A.stride(1), _bstride(A), ptr(B),
            B.stride(1), _bstride(B), float(beta), ptr(out), out.stride(1), out.stride(0), nb,
            stream_handle())
    call("sslab_dgemm_strided_batched", *args)
    if _rec.ops is not None:
        _rec.ops.append(("sslab_dgemm_strided_batched", args, (A, B, out)))
    return out


class _GTerm(ctypes.Structure):
    _fields_ = [("A", ctypes.c_void_p), ("B", ctypes.c_void_p), ("strideA", ctypes.c_longlong),
                ("strideB", ctypes.c_longlong), ("lda", ctypes.c_int), ("ldb", ctypes.c_int),
                ("K", ctypes.c_int), ("transA", ctypes.c_int), ("transB", ctypes.c_int),
                ("alpha", ctypes.c_double)]


class _GDesc(ctypes.Structure):
    _fields_ = [("term", _GTerm * 3), ("nterms", ctypes.c_int), ("M", ctypes.c_int),
                ("N", ctypes.c_int), ("batch", ctypes.c_int), ("Cin", ctypes.c_void_p),
                ("strideCin", ctypes.c_longlong), ("ldcin", ctypes.c_int),
                ("C", ctypes.c_void_p), ("strideC", ctypes.c_longlong), ("ldc", ctypes.c_int),
                ("beta", ctypes.c_double)]


def gdesc(out, terms, cin=None, beta=0.0):
    """Descriptor of out_b = beta cin_b + sum_t alpha_t op(A_t,b) op(B_t,b) for
    sslab_dgemm_grouped; terms = [(A, B, alpha[, ta, tb]), ...] (up to 3, (nb, ., .)
    tensors, a batch of 1 broadcasts).  Returns (descriptor, tensors to keep alive)."""
    if not 1 <= len(terms) <= 3:
        raise ValueError("gdesc: 1..3 terms")
    _check_f64("gdesc", out, cin, *[x for t in terms for x in t[:2]])
    if out.stride(2) != 1:
        raise ValueError("gdesc: out needs unit column stride")
    d = _GDesc()
    keep = [out]
    nb, M, N = out.shape
    for i, t in enumerate(terms):
        A, B, alpha = t[0], t[1], t[2]
        ta = bool(t[3]) if len(t) > 3 else False
        tb = bool(t[4]) if len(t) > 4 else False
        if A.stride(2) != 1:
            _recording_copy()
            A = A.contiguous()
        if B.stride(2) != 1:
            _recording_copy()
            B = B.contiguous()
        Ma = A.shape[2] if ta else A.shape[1]
        K = A.shape[1] if ta else A.shape[2]
        K2 = B.shape[2] if tb else B.shape[1]
        Nb = B.shape[1] if tb else B.shape[2]
        if Ma != M or Nb != N or K != K2 or A.shape[0] not in (1, nb) or B.shape[0] not in (1, nb):
            raise ValueError(f"gdesc: term {i} does not match the output")
        keep += [A, B]
        d.term[i] = _GTerm(ptr(A), ptr(B), _bstride(A), _bstride(B), A.stride(1), B.stride(1),
                           K, int(ta), int(tb), float(alpha))
    d.nterms, d.M, d.N, d.batch = len(terms), M, N, nb
    d.C, d.strideC, d.ldc = ptr(out), out.stride(0), out.stride(1)
    if cin is not None and beta != 0.0:
        if tuple(cin.shape) != (nb, M, N) or cin.stride(2) != 1:
            raise ValueError("gdesc: cin must match the output")
        keep.append(cin)
        d.Cin, d.strideCin, d.ldcin, d.beta = ptr(cin), cin.stride(0), cin.stride(1), float(beta)
    else:
        d.Cin, d.strideCin, d.ldcin, d.beta = None, 0, N, 0.0
    return d, keep


def ggemm(*descs, stream=None):
    """Launch descriptors made by gdesc (independent outputs) as one grouped product."""
    arr = (_GDesc * len(descs))(*[d for d, _ in descs])
    args = (ctypes.addressof(arr), len(descs), stream_handle(stream))
    call("sslab_dgemm_grouped", *args)
    if _rec.ops is not None:
        _rec.ops.append(("sslab_dgemm_grouped", args, (arr, [k for _, k in descs])))


def bqr(A, want_q=True, pivot=False):
    """Batched Householder QR of (nb, m, n): returns (Q (nb, m, m) or None, R (nb, m, n),
    perm (nb, n) or None) with A[:, :, perm] = Q R."""
    _check_f64("bqr", A)
    R = A.contiguous().clone()
    nb, m, n = R.shape
    Q = torch.empty(nb, m, m, dtype=F64, device=A.device) if want_q else None
    perm = torch.empty(nb, n, dtype=I32, device=A.device) if pivot else None
    call("sslab_dgeqrf_batched", m, n, ptr(R), n, m * n, ptr(Q) if want_q else None, m, m * m,
         ptr(perm) if pivot else None, n, nb, stream_handle())
    return Q, R, perm


def btri_inv(R):
    """Inverse of a batch of upper-triangular (nb, n, n) blocks."""
    _check_f64("btri_inv", R)
    nb, n, _ = R.shape
    R = R.contiguous()
    X = torch.eye(n, dtype=F64, device=R.device).expand(nb, n, n).contiguous()
    if n:
        call("sslab_dtrsm_upper_batched", n, n, ptr(R), n, n * n, ptr(X), n, n * n, nb,
             stream_handle())
    return X


def _t(X):
    return X.transpose(1, 2)


# ---------------------------------------------------------------------------

class DeviceHBS:
    """HBS matrix with device-resident factors (lists ordered leaf level first)."""

    def __init__(self, Umats, Dmats, Vmats, fac=2, perm=None):
        _check_f64("DeviceHBS", *Umats, *Dmats, *Vmats)
        self.Umats = list(Umats)
        self.Dmats = list(Dmats)
        self.Vmats = list(Vmats)
        self.fac = fac
        self.device = self.Dmats[0].device
        N = self.Dmats[0].shape[0] * self.Dmats[0].shape[1]
        self.N = N
        self.perm = device_perm(perm, self.device)
        self.Nbvec = [D.shape[0] for D in self.Dmats]
        self._ulv = {}
        self._plans = {}

    # -- construction from sketches (HBSnew.py:147-210) --------------------------------
    @staticmethod
    def from_sketches(Om, Psi, Y, Z, Nb, L, rk, fac=2, perm=None):
        """Om, Psi (N x s) test matrices and Y = A Om, Z = A^T Psi, all in leaf-permuted
        order, on the device."""
        _check_f64("DeviceHBS.from_sketches", Om, Psi, Y, Z)
        N, s = Om.shape
        nl = N // Nb
        nb, n = Nb, nl
        Oml, Psl, Yl, Zl = Om, Psi, Y, Z
        Umats, Vmats, Dmats = [], [], []
        U = V = D = None
        for lvl in range(L - 1, -1, -1):
            if lvl < L - 1:  # sketches of the next level: two grouped launches
                O3 = Oml.view(nb, n, s)
                P3 = Psl.view(nb, n, s)
                kU, kV = U.shape[2], V.shape[2]
                dev = O3.device
                Ynew = torch.empty(nb, n, s, dtype=F64, device=dev)
                Znew = torch.empty(nb, n, s, dtype=F64, device=dev)
                On = torch.empty(nb, kV, s, dtype=F64, device=dev)
                Pn = torch.empty(nb, kU, s, dtype=F64, device=dev)
                ggemm(gdesc(Ynew, [(D, O3, -1.0)], cin=Yl.view(nb, n, s), beta=1.0),
                      gdesc(Znew, [(D, P3, -1.0, True)], cin=Zl.view(nb, n, s), beta=1.0),
                      gdesc(On, [(V, O3, 1.0, True)]), gdesc(Pn, [(U, P3, 1.0, True)]))
                Yn = torch.empty(nb, kU, s, dtype=F64, device=dev)
                Zn = torch.empty(nb, kV, s, dtype=F64, device=dev)
                ggemm(gdesc(Yn, [(U, Ynew, 1.0, True)]), gdesc(Zn, [(V, Znew, 1.0, True)]))
                Yl, Zl = Yn.reshape(-1, s), Zn.reshape(-1, s)
                Oml, Psl = On.reshape(-1, s), Pn.reshape(-1, s)
                nb //= fac
                n = Oml.shape[0] // nb
            rkm = min(rk, nl * fac ** (L - 1 - lvl))
            O3 = Oml.reshape(nb, n, s)
            P3 = Psl.reshape(nb, n, s)
            Y3 = Yl.reshape(nb, n, s)
            Z3 = Zl.reshape(nb, n, s)
            if lvl > 0:
                U, V, pinvO, pinvP = _basis_and_pinv_pair(O3, P3, Y3, Z3, rkm)
                # T1 = (I - U U^T) Y pinv(Omega), T2 = (I - V V^T) Z pinv(Psi): each step of
                # the two sides in one grouped launch
                dev = Y3.device
                T1 = torch.empty(nb, n, n, dtype=F64, device=dev)
                T2 = torch.empty(nb, n, n, dtype=F64, device=dev)
                ggemm(gdesc(T1, [(Y3, pinvO, 1.0)]), gdesc(T2, [(Z3, pinvP, 1.0)]))
                W1 = torch.empty(nb, U.shape[2], n, dtype=F64, device=dev)
                W2 = torch.empty(nb, V.shape[2], n, dtype=F64, device=dev)
                ggemm(gdesc(W1, [(U, T1, 1.0, True)]), gdesc(W2, [(V, T2, 1.0, True)]))
                ggemm(gdesc(T1, [(U, W1, -1.0)], cin=T1, beta=1.0),
                      gdesc(T2, [(V, W2, -1.0)], cin=T2, beta=1.0))
                D = bgemm(U, bgemm(U, T2, ta=True, tb=True), T1, 1.0, 1.0)
                Umats.append(U)
                Vmats.append(V)
                Dmats.append(D)
            else:
                _, pinvO = _basis_and_pinv(O3, None, 0)
                Dmats.append(bgemm(Y3, pinvO))
        return DeviceHBS(Umats, Dmats, Vmats, fac, perm)

    # -- apply (HBSnew.py:303-388) ----------------------------------------------------
    def apply(self, x: torch.Tensor, mode="N") -> torch.Tensor:
        one_d = x.dim() == 1
        X = x.reshape(x.shape[0], -1).to(self.device, F64)
        u = self._planned("A", mode, X, lambda inp: self._apply_core(inp, mode))
        return u[:, 0] if one_d else u

    # -- launch plans -------------------------------------------------------------------
    def _planned(self, kind, mode, B, core):
        """core(inp) on B (N x r, leaf order restored on return) through a launch plan per
        (kind, mode, r, stream): the first call runs core on plan-owned buffers and records
        its GEMM launches (raw arguments); later calls gather B into the plan's input and
        replay them, so the per-call host work is the launches alone.  The result is always
        a new tensor.  Plans are not built or used while the stream is being captured, and
        a core whose launches needed operand copies is never replayed."""
        r = B.shape[1]
        if not _PLANS or r == 0 or torch.cuda.is_current_stream_capturing():
            inp = B[self.perm] if self.perm is not None else B.contiguous()
            return self._unpermute(core(inp), fresh=False)
        key = (kind, mode, r, stream_handle())
        plan = self._plans.get(key)
        if plan is not None:
            inp, out, ops = plan
            self._gather_into(inp, B)
            for name, args, _ in ops:
                call(name, *args)
            return self._unpermute(out, fresh=True)
        inp = torch.empty(self.N, r, dtype=F64, device=self.device)
        self._gather_into(inp, B)
        _rec.ops, _rec.bad = [], False
        try:
            out = core(inp)
            ops, bad = _rec.ops, _rec.bad
        finally:
            _rec.ops, _rec.bad = None, False
        if not bad:
            if len(self._plans) >= 16:  # many widths or streams: drop the oldest plan
                self._plans.pop(next(iter(self._plans)))
            self._plans[key] = (inp, out, ops)
        return self._unpermute(out, fresh=True)

    def _gather_into(self, inp, B):
        if self.perm is not None:
            torch.index_select(B, 0, self.perm, out=inp)
        else:
            inp.copy_(B)

    def _unpermute(self, y, fresh):
        if self.perm is not None:
            w = torch.empty_like(y)
            w[self.perm] = y
            return w
        return y.clone() if fresh else y

    def _apply_core(self, X, mode):
        r = X.shape[1]
        In, Out = (self.Vmats, self.Umats) if mode == "N" else (self.Umats, self.Vmats)
        tD = mode != "N"
        VV = [X.contiguous()]
        for lvl, M in enumerate(In):
            nb, n, k = M.shape
            VV.append(bgemm(M, VV[lvl].view(nb, n, r), ta=True).reshape(nb * k, r))
        Dr = self.Dmats[-1]
        u = bgemm(Dr, VV[-1].view(Dr.shape[0], -1, r), ta=tD).reshape(-1, r)
        for lvl in range(len(Out) - 1, -1, -1):
            M = Out[lvl]
            nb, n, k = M.shape
            out = torch.empty(nb, n, r, dtype=F64, device=self.device)
            ggemm(gdesc(out, [(M, u.view(nb, k, r), 1.0),
                              (self.Dmats[lvl], VV[lvl].view(nb, n, r), 1.0, tD)]))
            u = out.reshape(nb * n, r)
        return u

    def dense(self):
        return self.apply(torch.eye(self.N, dtype=F64, device=self.device))

    @property
    def nbytes(self):
        return 8 * sum(M.numel() for M in self.Umats + self.Vmats + self.Dmats)

    def transpose(self):
        return DeviceHBS(self.Vmats, [_t(D).contiguous() for D in self.Dmats], self.Umats,
                         self.fac, self.perm)

    def split(self, K, perm=None):
        """The K matrices of a stacked DeviceHBS (built from K operators' sketches stacked
        along every level's batch, so matrix k owns boxes [k nb_l / K, (k+1) nb_l / K) of
        each level, root k included), with their ULV factors if any were computed."""
        def part(M, k):
            n = M.shape[0] // K
            return M[k * n:(k + 1) * n]

        out = []
        for k in range(K):
            h = DeviceHBS([part(U, k) for U in self.Umats], [part(D, k) for D in self.Dmats],
                          [part(V, k) for V in self.Vmats], self.fac, perm)
            for mode, levels in self._ulv.items():
                h._ulv[mode] = [{key: (part(v, k) if isinstance(v, torch.Tensor)
                                       else (v // K if key == "nb" else v))
                                 for key, v in lev.items()} for lev in levels]
            out.append(h)
        return out

    # -- ULV (ULVsparse.py:231-358) -----------------------------------------------------
    def ulv_factor(self, mode="N"):
        if mode in self._ulv:
            return self._ulv[mode]
        if mode == "T":
            f = self.transpose().ulv_factor("N")
            self._ulv["T"] = f
            return f
        fac = self.fac
        levels = []
        Uhat = R22 = None
        nlev = len(self.Dmats)
        # Q of every level's V depends on V alone: one QR launch per distinct box shape
        # (leaf level and upper levels) instead of one per level in the sequential sweep
        QVs = [None] * (nlev - 1)
        groups: dict = {}
        for lvl in range(nlev - 1):
            groups.setdefault(tuple(self.Vmats[lvl].shape[1:]), []).append(lvl)
        for lvls in groups.values():
            Qs, _, _ = bqr(torch.cat([self.Vmats[lvl] for lvl in lvls]))
            off = 0
            for lvl in lvls:
                nbl = self.Vmats[lvl].shape[0]
                QVs[lvl] = Qs[off:off + nbl]
                off += nbl
        for lvl in range(nlev):
            D = self.Dmats[lvl]
            if lvl == 0:
                Rhat = D
            else:
                nbp, nf, _ = D.shape
                k = nf // fac
                Rhat = bgemm(Uhat, D.reshape(nbp * fac, k, nf)).reshape(nbp, nf, nf)
                call("sslab_add_diag_blocks", k, fac, nbp, ptr(R22.contiguous()), ptr(Rhat), nf,
                     nf * nf, stream_handle())
            if lvl == nlev - 1:
                Q, R, _ = bqr(Rhat)
                n = R.shape[1]
                G = bgemm(btri_inv(R), Q, tb=True)       # R^{-1} Q^T
                levels.append(dict(root=True, G=G, D=D))
                break
            V = self.Vmats[lvl]
            U = self.Umats[lvl]
            nb, n, k = V.shape
            QV = QVs[lvl]
            W = torch.cat([QV[:, :, k:], V], dim=2).contiguous()
            Q, R, _ = bqr(bgemm(Rhat, W))
            if lvl == 0:
                Uc = U
            else:  # blkdiag(Ud of the fac children) U: one product per child row block
                kc = Uhat.shape[2]
                Uc = bgemm(Uhat, U.reshape(nb * fac, kc, k)).reshape(nb, n, k)
            if n > k:  # Uu = Q_1^T Uc and Ud = Q_2^T Uc in one grouped launch
                Uu = torch.empty(nb, n - k, Uc.shape[2], dtype=F64, device=Uc.device)
                Ud = torch.empty(nb, k, Uc.shape[2], dtype=F64, device=Uc.device)
                ggemm(gdesc(Uu, [(Q[:, :, :n - k], Uc, 1.0, True)]),
                      gdesc(Ud, [(Q[:, :, n - k:], Uc, 1.0, True)]))
            else:
                Uu, Ud = None, bgemm(Q[:, :, n - k:], Uc, ta=True)
            R22 = R[:, n - k:, n - k:].contiguous()
            lev = dict(root=False, nb=nb, n=n, k=k, Q=Q, W=W, Uu=Uu,
                       R12=R[:, :n - k, n - k:].contiguous(),
                       R11inv=btri_inv(R[:, :n - k, :n - k].contiguous()), U=U, D=self.Dmats[lvl])
            levels.append(lev)
            Uhat = Ud
        self._ulv[mode] = levels
        return levels

    def ulv_factor_pair(self):
        """The ULV factors of A and A^T from one stacked factorization: the transposed
        HBS (V, D^T, U) is appended to this one along every level's batch (block p of
        the second half has its children at 2p, 2p+1 of the next level's second half,
        as in the first), so each level's QRs and products run once at twice the batch
        instead of twice in sequence.  Caches both modes; returns the mode-N levels."""
        if "N" in self._ulv and "T" in self._ulv:
            return self._ulv["N"]
        if len(self.Dmats) < 2:
            self.ulv_factor("T")
            return self.ulv_factor("N")
        st = DeviceHBS([torch.cat([U, V]) for U, V in zip(self.Umats, self.Vmats)],
                       [torch.cat([D, _t(D)]) for D in self.Dmats],
                       [torch.cat([V, U]) for U, V in zip(self.Umats, self.Vmats)], self.fac)
        both = st.ulv_factor("N")
        halves = ([], [])
        for lev in both:
            for h, out in enumerate(halves):
                d = {}
                for key, val in lev.items():
                    if isinstance(val, torch.Tensor):
                        nb2 = val.shape[0]
                        d[key] = val[: nb2 // 2] if h == 0 else val[nb2 // 2:]
                    elif key == "nb":
                        d[key] = val // 2
                    else:
                        d[key] = val
                out.append(d)
        self._ulv["N"], self._ulv["T"] = halves
        return self._ulv["N"]

    def solve(self, b: torch.Tensor, mode="N") -> torch.Tensor:
        """x with A x = b (mode N) or A^T x = b (mode T), b of shape (N,) or (N, r)."""
        levels = self.ulv_factor(mode)  # before any plan records: the factor is not replayed
        one_d = b.dim() == 1
        B = b.reshape(b.shape[0], -1).to(self.device, F64)
        y = self._planned("S", mode, B, lambda inp: self._solve_core(levels, inp))
        return y[:, 0] if one_d else y

    def _solve_core(self, levels, B):
        r = B.shape[1]
        cur = B.contiguous()
        tops = []
        for lev in levels[:-1]:
            nb, n, k, Q = lev["nb"], lev["n"], lev["k"], lev["Q"]
            c3 = cur.view(nb, n, r)
            nxt = torch.empty(nb, k, r, dtype=F64, device=self.device)
            if n > k:  # both halves of Q^T c in one launch
                top = torch.empty(nb, n - k, r, dtype=F64, device=self.device)
                ggemm(gdesc(top, [(Q[:, :, :n - k], c3, 1.0, True)]),
                      gdesc(nxt, [(Q[:, :, n - k:], c3, 1.0, True)]))
                tops.append(top)
            else:
                bgemm(Q[:, :, n - k:], c3, nxt, ta=True)
                tops.append(None)
            cur = nxt.reshape(nb * k, r)
        root = levels[-1]
        y = bgemm(root["G"], cur.view(1, -1, r))
        v = bgemm(root["D"], y)
        y = y.reshape(-1, r)
        v = v.reshape(-1, r)
        for lvl in range(len(levels) - 2, -1, -1):
            lev = levels[lvl]
            nb, n, k = lev["nb"], lev["n"], lev["k"]
            y2 = y.view(nb, k, r)
            vb = v.view(nb, k, r)
            x = torch.empty(nb, n, r, dtype=F64, device=self.device)
            if n > k:  # summed terms: one launch per dependent product
                rhs0 = torch.empty_like(tops[lvl])
                ggemm(gdesc(rhs0, [(lev["Uu"], vb, -1.0), (lev["R12"], y2, -1.0)],
                            cin=tops[lvl], beta=1.0))
                y1 = bgemm(lev["R11inv"], rhs0)
                ggemm(gdesc(x, [(lev["W"][:, :, :n - k], y1, 1.0), (lev["W"][:, :, n - k:], y2, 1.0)]))
            else:
                bgemm(lev["W"], y2, x)
            vn = torch.empty(nb, n, r, dtype=F64, device=self.device)
            ggemm(gdesc(vn, [(lev["U"], vb, 1.0), (lev["D"], x, 1.0)]))
            v = vn.reshape(nb * n, r)
            y = x.reshape(nb * n, r)
        return y


def _basis_and_pinv_pair(O3, P3, Y3, Z3, rk):
    """_basis_and_pinv of the column side (O3, Y3) and the row side (P3, Z3) of one level
    in stacked launches of twice the batch: the two sides are independent, and each
    batched QR runs one workgroup per box, so the stacked launch costs about what one
    side's does."""
    nb, n, s = O3.shape
    if s <= n:
        raise ValueError(f"HBS compression needs more samples than box rows ({s} <= {n})")
    OP = torch.empty(2 * nb, s, n, dtype=F64, device=O3.device)
    OP[:nb].copy_(O3.transpose(1, 2))
    OP[nb:].copy_(P3.transpose(1, 2))
    Q, R, _ = bqr(OP)                             # (2nb, s, s), R (2nb, s, n)
    Rinv = btri_inv(R[:, :n, :n].contiguous())
    pinv = bgemm(Q[:, :, :n], Rinv, tb=True)      # (2nb, s, n)
    YZ = torch.cat([Y3, Z3])
    M = bgemm(YZ, Q[:, :, n:])                    # (2nb, n, s - n)
    QM, _, _ = bqr(M, pivot=True)
    B = QM[:, :, :rk]
    return B[:nb].contiguous(), B[nb:].contiguous(), pinv[:nb], pinv[nb:]


def _basis_and_pinv(Om3, Y3, rk):
    """Per box: complete QR of Om_i^T = Q R; the basis of Y_i Q[:, n:] (Y_i restricted
    to the null space of Om_i, which removes the diagonal-block term, as in
    matAssembly/HBS/HBStorch.py:73-82) from a column-pivoted QR (its first rk columns);
    pinv(Om_i) = Q[:, :n] R^{-T}.  Needs s > n samples."""
    nb, n, s = Om3.shape
    if s <= n:
        raise ValueError(f"HBS compression needs more samples than box rows ({s} <= {n})")
    Q, R, _ = bqr(_t(Om3))                        # (nb, s, s), R (nb, s, n)
    Rinv = btri_inv(R[:, :n, :n].contiguous())
    pinv = bgemm(Q[:, :, :n], Rinv, tb=True)      # (nb, s, n)
    if Y3 is None:
        return None, pinv
    M = bgemm(Y3, Q[:, :, n:])                    # (nb, n, s - n)
    QM, _, _ = bqr(M, pivot=True)
    return QM[:, :, :rk].contiguous(), pinv
