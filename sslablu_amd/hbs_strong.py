"""Strong-admissibility HBS matrices on the MI355X: compression from sketches and apply.

Device engine behind the mirrors of matAssembly/HBS/HBStorch_strong.py (HBSMAT) and
matAssembly/HBS/HBSnew_strong.py (HBSStrong).  Format (HBStorch_strong.py:1232-1294):
per compressed level, A = U B V^T + E with U, V (nb, b, r) orthonormal cluster bases, E
the near-field error term on the pairs (i, j), j in NF(i) = {i} + adj(i), and B the
operator of the next level (children regrouped by fac); a dense base closes the
recursion once every box's near field is the whole level.

Construction per level (HBStorch_strong.py:687-797, 1033-1107), batched over the boxes
that share a near-field size k (one launch per group and step):
  * S_i = [Om_j ; j in NF(i)] (k b x s) gathered from the test-matrix blocks;
  * G_i = S_i S_i^T and W_i = Y_i S_i^T (MFMA GEMMs), G_i factored by the batched LU;
    P_i = W_i G_i^{-1} = Y_i pinv(S_i) (S_i has full row rank: Gaussian blocks, s >= k b);
  * the far-field part M_i = Y_i - P_i S_i = Y_i (I - pi_i), pi_i the projector on the
    row space of S_i; U_i = the first r columns of a column-pivoted QR of the b x b Gram
    matrix M_i M_i^T (the reference takes the leading left singular vectors of M_i; all
    span the far field when it has rank <= r, and the Gram form keeps the pivoted QR at
    b x b, inside one workgroup's LDS, instead of b x s);
  * term1 tiles (I - U_i U_i^T) P_i = (I - U_i U_i^T) A_ij (j in NF(i)); the same on the
    Psi / Z side gives V_j and (I - V_j V_j^T) A_ij^T (i in NF(j)); E_ij = term1 +
    U_i U_i^T A_ij (I - V_j V_j^T);
  * next level: Y <- U^T (Y - E Om), Z <- V^T (Z - E^T Psi), Om <- V^T Om,
    Psi <- U^T Psi;
  * base: B = Y Om^T (Om Om^T)^{-1} over the whole level.
The reference factors S_i^T by QR (_factor_nf :261-287); here the Gram matrix of the
Gaussian stack is factored instead (condition number kappa(S)^2, a few hundred for the
sample counts used), which keeps every step on the GEMM and LU kernels.

E is stored per box row as (nb, b, g b) with g the largest near-field size (slots past a
box's own k hold zeros), together with its transpose layout for mode T, so a near-field
apply is one gather of g source blocks per box and one batched GEMM.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import call, ptr, stream_handle
from .hbs import _check_f64, bgemm, bqr, device_perm, gdesc, ggemm

F64 = torch.float64
I32 = torch.int32


def _lu_solve(G, B):
    """X = G_b^{-1} B_b for (nb, n, n) G (overwritten by its LU factors) and (nb, n, r) B."""
    nb, n, _ = G.shape
    r = B.shape[2]
    piv = torch.empty(nb, n, dtype=I32, device=G.device)
    info = torch.zeros(nb, dtype=I32, device=G.device)
    s = stream_handle()
    call("sslab_dgetrf_batched", n, ptr(G), n, n * n, ptr(piv), n, ptr(info), nb, s)
    B = B.contiguous()
    X = torch.empty_like(B)
    perm = torch.empty(nb, n, dtype=I32, device=G.device)
    call("sslab_dgetrs_batched", n, r, ptr(G), n, n * n, ptr(piv), n, ptr(B), r, n * r,
         ptr(X), r, n * r, ptr(perm), nb, s)
    return X, info


def _has_far(adj_l):
    n = len(adj_l)
    return any(len(set(adj_l[i]) | {i}) < n for i in range(n))


class _NearField:
    """Sorted near fields of one level, their padded slot layout and the groups of equal
    size (NearField, HBStorch_strong.py:199-245)."""

    def __init__(self, adj_l, device):
        n = len(adj_l)
        for i in range(n):
            if i in adj_l[i] or any(i not in adj_l[j] for j in adj_l[i]):
                raise ValueError("adjacency must be symmetric and exclude the box itself")
        self.nf = [sorted(set(adj_l[i]) | {i}) for i in range(n)]
        self.n = n
        self.g = max(len(f) for f in self.nf)
        g = self.g
        slot = [{j: a for a, j in enumerate(f)} for f in self.nf]
        # source box per slot (pad: the extra zero block n) and, for the transposed
        # layout, the flat tile index j * g + (slot of i in NF(j)) (pad: n * g)
        src = np.full((n, g), n, dtype=np.int64)
        back = np.full((n, g), n * g, dtype=np.int64)
        for i, f in enumerate(self.nf):
            for a, j in enumerate(f):
                src[i, a] = j
                back[i, a] = j * g + slot[j][i]
        self.src = torch.as_tensor(src.ravel(), device=device)
        self.back = torch.as_tensor(back.ravel(), device=device)
        groups = {}
        for i, f in enumerate(self.nf):
            groups.setdefault(len(f), []).append(i)
        self.groups = [(k, torch.as_tensor(r, device=device),
                        torch.as_tensor(np.array([self.nf[i] for i in r]).ravel(), device=device))
                       for k, r in sorted(groups.items())]


def _gather_blocks(X3, src):
    """(n, b, c) blocks -> (n, g b, c): row block a of box i is X3[src[i, a]] (zero for
    src = n)."""
    n, b, c = X3.shape
    Xz = torch.cat([X3, X3.new_zeros(1, b, c)])
    g = src.numel() // n
    return Xz.index_select(0, src).view(n, g * b, c)


def _tile_transpose_gather(T, idx, n, g, b):
    """T (n, b, g b) of tiles T[i][:, a] -> (n, b, g b) whose slot a of row j holds
    (tile at flat index idx[j, a])^T (zero for idx = n g)."""
    Tt = T.view(n, b, g, b).permute(0, 2, 3, 1).reshape(n * g, b, b)
    Tt = torch.cat([Tt, Tt.new_zeros(1, b, b)])
    return Tt.index_select(0, idx).view(n, g, b, b).permute(0, 2, 1, 3).reshape(n, b, g * b)


class DeviceHBSStrong:
    """Strong-admissibility HBS matrix with device-resident factors (leaf level first)."""

    def __init__(self, levels, B, fac, perm, N):
        self.levels = levels
        self.B = B
        self.fac = fac
        self.N = N
        self.device = B.device
        self.perm = perm

    # -- construction (HBStorch_strong.py:1033-1107) ------------------------------------
    @staticmethod
    def from_sketches(Om, Psi, Y, Z, adj, rk, fac=2, perm=None):
        """Om, Psi (N x s) test matrices and Y = A Om, Z = A^T Psi, in leaf order, on the
        device; adj[l] the neighbour sets of the boxes of tree level l (level-local ranks,
        children of rank P are fac P .. fac P + fac - 1), leaf level last."""
        _check_f64("DeviceHBSStrong.from_sketches", Om, Psi, Y, Z)
        N, s = Om.shape
        if s > 1024:
            raise ValueError(f"strong HBS: at most 1024 samples ({s})")
        dev = Om.device
        lvl = len(adj) - 1
        nb = len(adj[lvl])
        if N % nb:
            raise ValueError(f"strong HBS: {N} rows do not split into {nb} equal leaves")
        b = N // nb
        O3, P3 = Om.reshape(nb, b, s), Psi.reshape(nb, b, s)
        Y3, Z3 = Y.reshape(nb, b, s), Z.reshape(nb, b, s)
        levels = []
        while _has_far(adj[lvl]):
            nb = len(adj[lvl])
            if nb % fac or O3.shape[0] != nb:
                raise ValueError("strong HBS: the tree levels do not match the blocks")
            nfo = _NearField(adj[lvl], dev)
            g = nfo.g
            rkm = min(rk, b)
            if g * b > s:
                raise ValueError(f"strong HBS: {s} samples < near-field rows {g * b} at "
                                 f"level {lvl}")
            U = torch.empty(nb, b, rkm, dtype=F64, device=dev)
            V = torch.empty_like(U)
            X1 = torch.zeros(nb, b, g * b, dtype=F64, device=dev)
            X2 = torch.zeros_like(X1)
            for S3, K3, Bout, Xout in ((O3, Y3, U, X1), (P3, Z3, V, X2)):
                for k, ranks, nf_idx in nfo.groups:
                    G = ranks.numel()
                    Snf = S3.index_select(0, nf_idx).view(G, k * b, s)
                    Kg = K3.index_select(0, ranks)
                    Gm = bgemm(Snf, Snf, tb=True)                       # (G, kb, kb)
                    Wt = bgemm(Snf, Kg, tb=True)                        # (G, kb, b)
                    Pt, _ = _lu_solve(Gm, Wt)                           # P^T = G^{-1} W^T
                    M = bgemm(Pt, Snf, out=Kg, alpha=-1.0, beta=1.0, ta=True)
                    Qm, _, _ = bqr(bgemm(M, M, tb=True), pivot=True)     # b x b Gram
                    Bb = Qm[:, :, :rkm].contiguous()                    # (G, b, r)
                    C = bgemm(Bb, Pt, ta=True, tb=True)                 # (G, r, kb)
                    P = Pt.transpose(1, 2).contiguous()
                    X = bgemm(Bb, C, out=P, alpha=-1.0, beta=1.0)       # (I - B B^T) P
                    Bout.index_copy_(0, ranks, Bb)
                    Xout[:, :, :k * b].index_copy_(0, ranks, X)         # later slots stay 0
            # E = term1 + U U^T term2, term2 tiles from the transposed Psi-side tiles
            R2 = _tile_transpose_gather(X2, nfo.back, nb, g, b)
            E = bgemm(U, bgemm(U, R2, ta=True), out=X1, beta=1.0)
            Et = _tile_transpose_gather(E, nfo.back, nb, g, b)
            # reduction to the next level
            Yn = bgemm(E, _gather_blocks(O3, nfo.src), out=Y3.clone(), alpha=-1.0, beta=1.0)
            Zn = bgemm(Et, _gather_blocks(P3, nfo.src), out=Z3.clone(), alpha=-1.0, beta=1.0)
            r2 = U.shape[2]
            nxt = nb // fac
            Y3n = torch.empty(nb, r2, s, dtype=F64, device=dev)
            Z3n, O3n, P3n = torch.empty_like(Y3n), torch.empty_like(Y3n), torch.empty_like(Y3n)
            ggemm(gdesc(Y3n, [(U, Yn, 1.0, True)]), gdesc(Z3n, [(V, Zn, 1.0, True)]),
                  gdesc(O3n, [(V, O3, 1.0, True)]), gdesc(P3n, [(U, P3, 1.0, True)]))
            levels.append(dict(U=U, V=V, E=E, Et=Et, src=nfo.src, g=g, b=b, nb=nb))
            b = fac * r2
            O3, P3 = O3n.view(nxt, b, s), P3n.view(nxt, b, s)
            Y3, Z3 = Y3n.view(nxt, b, s), Z3n.view(nxt, b, s)
            lvl -= 1
            if lvl < 0:
                raise ValueError("strong HBS: the root level still has a far field")
        nbase = O3.shape[0] * b
        if nbase > s:
            raise ValueError(f"strong HBS: {s} samples < base rows {nbase}")
        Of, Yf = O3.reshape(1, nbase, s), Y3.reshape(1, nbase, s)
        Gm = bgemm(Of, Of, tb=True)
        Bt, _ = _lu_solve(Gm, bgemm(Of, Yf, tb=True))                   # B^T = G^{-1} Om Y^T
        B = Bt[0].t().contiguous()
        return DeviceHBSStrong(levels, B, fac, perm, N)

    # -- apply (HBStorch_strong.py:1232-1294) -------------------------------------------
    def apply(self, x: torch.Tensor, mode="N") -> torch.Tensor:
        """A x (mode 'N') or A^T x ('T') for x (N,) or (N, r) in the original order."""
        _check_f64("DeviceHBSStrong.apply", x)
        one = x.dim() == 1
        X = x.reshape(self.N, -1)
        if self.perm is not None:
            X = X.index_select(0, self.perm)
        X = X.contiguous()
        c = X.shape[1]
        tr = mode == "T"
        down = []
        g = X
        for lv in self.levels:
            g3 = g.view(lv["nb"], lv["b"], c)
            down.append(g3)
            g = bgemm(lv["U" if tr else "V"], g3, ta=True).view(-1, c)
        w = bgemm(self.B[None], g[None], ta=tr)[0]
        for lv, g3 in zip(reversed(self.levels), reversed(down)):
            nb = lv["nb"]
            Bu = lv["V" if tr else "U"]
            near = _gather_blocks(g3, lv["src"])
            out = torch.empty(nb, lv["b"], c, dtype=F64, device=self.device)
            ggemm(gdesc(out, [(Bu, w.view(nb, -1, c), 1.0), (lv["Et" if tr else "E"], near, 1.0)]))
            w = out.view(-1, c)
        if self.perm is not None:
            y = torch.empty_like(w)
            y.index_copy_(0, self.perm, w)
            w = y
        return w.reshape(self.N) if one else w

    def dense(self):
        return self.apply(torch.eye(self.N, dtype=F64, device=self.device))

    @property
    def nbytes(self):
        tot = self.B.numel() * 8
        for lv in self.levels:
            tot += (lv["U"].numel() + lv["V"].numel() + lv["E"].numel()) * 8
        return tot

    @property
    def nlevels(self):
        return len(self.levels)


def strong_sample_count(adj, leaf, rk, p=20, fac=2):
    """Gaussian samples for a tree adjacency: kmax max(fac rk, leaf) + rk + p, kmax the
    largest near-field size over the levels (test/rkprop/thinSlab3D.py:400-408 for binary
    trees, fac = 2; a quad tree's coarse boxes hold 4 rk rows, HBSnew_strong.py:404)."""
    kmax = max(max((len(a) + 1 for a in al), default=1) for al in adj)
    return kmax * max(fac * rk, leaf) + rk + p


def perm_tensor(perm, device):
    return device_perm(perm, device)
