"""Oracle: HBS (hierarchically block-separable) compression, apply and ULV solve
(numpy FP64, CPU).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates
  * binary line tree and its leaf permutation -- matAssembly/HBS/slabTree.py:206-237
  * exactly-HBS random operators -- test/rkprop/HBSnewTest.py:15-73 (gen_random_HBS)
  * telescoping apply A = D + U (D' + U' (...) V'^T) V^T and its transpose
    -- matAssembly/HBS/HBSnew.py:303-388 (matmat / rmatmat)
  * black-box compression from sketches Y = A Om, Z = A^T Psi
    -- HBSnew.py:147-210 (constructHBS) with compute_UV (:20-28), construct_D (:48-62),
       block_solve_r (:11-19); sample count s = 4k + 2 nl + 5 (binary) / 6k + 4 nl + 5 (quad)
  * ULV factorization Q^T A W = R and solve -- matAssembly/HBS/ULVsparse.py:18-358
    (compute_QRW_sparse, compute_ULV, solve)

Layout ("reduced form", HBSnew.py:65-139): level l holds Nb_l boxes; Umats[l], Vmats[l]
are (Nb_l, n_l, k) stacks, Dmats[l] is (Nb_l, n_l, n_l); the last entry of Dmats is the
root block.  n_0 = leaf size, n_{l+1} = fac * k.

Parity pins (tests/test_oracle_hbs.py): compressing an exactly-HBS operator reproduces
its apply to ~1e-12 (the check of HBSnewTest.py:89-109); the ULV solve inverts the
dense HBS matrix (vs numpy.linalg.solve) in modes N and T; W_i is orthogonal and
Q_i^T D_i W_i is upper triangular (ULVsparse.py:10-16).
"""
from __future__ import annotations

import numpy as np


# ---------------------------------------------------------------------------
# tree

def binary_line_tree(coords: np.ndarray, min_leaf: int):
    """Midpoint bisection of the points (slabTree._build_binary_line, slabTree.py:214-237)
    until a box has <= min_leaf points.  Returns (perm_leaf, Nb, nlevels); the leaves
    must come out equally sized (the reduced form needs a uniform leaf size)."""
    coords = np.asarray(coords, dtype=np.float64).ravel()
    leaves = []

    def split(idx, lo, hi, depth):
        if len(idx) <= min_leaf:
            leaves.append((idx, depth))
            return
        mid = 0.5 * (lo + hi)
        v = coords[idx]
        split(idx[v <= mid + 1e-10], lo, mid, depth + 1)
        split(idx[v >= mid - 1e-10], mid, hi, depth + 1)

    split(np.arange(len(coords)), coords.min(), coords.max(), 0)
    sizes = {len(l[0]) for l in leaves}
    depths = {l[1] for l in leaves}
    if len(sizes) != 1 or len(depths) != 1:
        raise ValueError("binary tree leaves are not uniform")
    perm = np.concatenate([l[0] for l in leaves])
    return perm, len(leaves), depths.pop() + 1


# ---------------------------------------------------------------------------
# generator / apply

def _orth(rng, n, k):
    return np.linalg.qr(rng.standard_normal((n, k)))[0]


def gen_random_hbs(k, nl, L, fac=2, rng=None):
    """Exactly-HBS operator in reduced form (HBSnewTest.py:15-73): orthonormal U, V per
    box, D_i = G - U (U^T G V) V^T on every non-root level, dense random root."""
    rng = np.random.default_rng(0) if rng is None else rng
    Nb = fac ** (L - 1)
    Umats, Vmats, Dmats, Nbvec = [], [], [], [Nb]
    k0 = min(k, nl)
    n = nl
    kk = k0
    for lvl in range(L):
        if lvl > 0:
            Nb //= fac
            Nbvec.append(Nb)
            n = fac * (k0 if lvl == 1 else k)
            kk = k
        if lvl < L - 1:
            U = np.stack([_orth(rng, n, kk) for _ in range(Nb)])
            V = np.stack([_orth(rng, n, kk) for _ in range(Nb)])
            G = rng.standard_normal((Nb, n, n))
            D = G - U @ (np.transpose(U, (0, 2, 1)) @ G @ V) @ np.transpose(V, (0, 2, 1))
            Umats.append(U)
            Vmats.append(V)
            Dmats.append(D)
        else:
            Dmats.append(rng.standard_normal((Nb, n, n)))
    return Umats, Dmats, Vmats, Nbvec


def _bd(Ms, X):
    """Block-diagonal product: X split into len(Ms) row blocks (block_mult, HBSnew.py:32-47)."""
    nb, n, k = Ms.shape
    return np.einsum("bij,bjr->bir", Ms, X.reshape(nb, k, -1)).reshape(nb * n, -1)


def _bdT(Ms, X):
    nb, n, k = Ms.shape
    return np.einsum("bji,bjr->bir", Ms, X.reshape(nb, n, -1)).reshape(nb * k, -1)


def hbs_apply(Umats, Dmats, Vmats, v, mode="N", perm=None):
    """A v (mode N) or A^T v (mode T) for the telescoping representation
    (HBSnew.py:303-388).  perm: leaf permutation (x_perm = x[perm])."""
    v = np.asarray(v, dtype=np.float64)
    one_d = v.ndim == 1
    if one_d:
        v = v[:, None]
    if perm is not None:
        v = v[perm]
    In, Out = (Vmats, Umats) if mode == "N" else (Umats, Vmats)
    tr = (lambda M: M) if mode == "N" else (lambda M: np.transpose(M, (0, 2, 1)))
    VV = [v]
    for lvl in range(len(In)):
        VV.append(_bdT(In[lvl], VV[lvl]))
    u = _bd(tr(Dmats[-1]), VV[-1])
    for lvl in range(len(Out) - 1, -1, -1):
        u = _bd(Out[lvl], u) + _bd(tr(Dmats[lvl]), VV[lvl])
    if perm is not None:
        w = np.empty_like(u)
        w[perm] = u
        u = w
    return u[:, 0] if one_d else u


def hbs_dense(Umats, Dmats, Vmats):
    N = Dmats[0].shape[0] * Dmats[0].shape[1]
    return hbs_apply(Umats, Dmats, Vmats, np.eye(N))


# ---------------------------------------------------------------------------
# compression from sketches

def _compute_UV(Om, Y, rk):
    """Per box: U = leading rk left singular vectors of Y_i Q[:, -n:], Q from the
    complete QR of Om_i^T (HBSnew.py:20-28)."""
    nb, n, _ = Om.shape
    out = np.zeros((nb, Y.shape[1], rk))
    for i in range(nb):
        Q = np.linalg.qr(Om[i].T, mode="complete")[0]
        out[i] = np.linalg.svd(Y[i] @ Q[:, -n:])[0][:, :rk]
    return out


def _construct_D(U, V, Y, Z, Om, Psi):
    """D_i = (Y_i - U U^T Y_i) Om_i^+ + U U^T ((Z_i - V V^T Z_i) Psi_i^+)^T
    (HBSnew.py:48-62)."""
    nb, n, _ = U.shape
    D = np.zeros((nb, n, n))
    for i in range(nb):
        Ui, Vi = U[i], V[i]
        D[i] = (Y[i] - Ui @ (Ui.T @ Y[i])) @ np.linalg.pinv(Om[i]) + \
            Ui @ (Ui.T @ ((Z[i] - Vi @ (Vi.T @ Z[i])) @ np.linalg.pinv(Psi[i])).T)
    return D


def hbs_construct(matmat, rmatmat, N, Nb, L, rk, fac=2, perm=None, rng=None, s=None):
    """Black-box HBS compression (HBSnew.py:147-210).  matmat / rmatmat apply A and A^T
    to N x s blocks in the original ordering.  Returns (Umats, Dmats, Vmats, Nbvec, s)."""
    rng = np.random.default_rng(0) if rng is None else rng
    nl = N // Nb
    if s is None:
        s = 6 * rk + 4 * nl + 5 if fac == 4 else 4 * rk + 2 * nl + 5
    perm = np.arange(N) if perm is None else np.asarray(perm)
    Om = rng.standard_normal((N, s))
    Psi = rng.standard_normal((N, s))
    Omp = np.zeros_like(Om)
    Omp[perm] = Om
    Psip = np.zeros_like(Psi)
    Psip[perm] = Psi
    Y = np.asarray(matmat(Omp))[perm]
    Z = np.asarray(rmatmat(Psip))[perm]
    Umats, Vmats, Dmats, Nbvec = [], [], [], []
    nb = Nb
    n = nl
    Oml, Psl, Yl, Zl = Om, Psi, Y, Z
    U = V = D = None
    for lvl in range(L - 1, -1, -1):
        if lvl < L - 1:
            # peel this level off the sketches (HBSnew.py:173-183)
            Yl = _bdT(U, Yl - _bd(D, Oml))
            Zl = _bdT(V, Zl - _bd(np.transpose(D, (0, 2, 1)), Psl))
            Oml = _bdT(V, Oml)
            Psl = _bdT(U, Psl)
            nb //= fac
            n = Oml.shape[0] // nb
        rkm = min(rk, nl * fac ** (L - 1 - lvl))
        Om3 = Oml.reshape(nb, n, s)
        Ps3 = Psl.reshape(nb, n, s)
        Y3 = Yl.reshape(nb, n, s)
        Z3 = Zl.reshape(nb, n, s)
        if lvl > 0:
            U = _compute_UV(Om3, Y3, rkm)
            V = _compute_UV(Ps3, Z3, rkm)
            D = _construct_D(U, V, Y3, Z3, Om3, Ps3)
            Umats.append(U)
            Vmats.append(V)
            Dmats.append(D)
        else:
            Dmats.append(np.stack([Y3[i] @ np.linalg.pinv(Om3[i], rcond=1e-15)
                                   for i in range(nb)]))
        Nbvec.append(nb)
    return Umats, Dmats, Vmats, Nbvec, s


# ---------------------------------------------------------------------------
# ULV

def _blockdiag_children(X, fac):
    """(nb_child, a, b) -> (nb_child/fac, fac*a, fac*b) block diagonals of fac children."""
    nc, a, b = X.shape
    out = np.zeros((nc // fac, fac * a, fac * b))
    for p in range(nc // fac):
        for j in range(fac):
            out[p, j * a:(j + 1) * a, j * b:(j + 1) * b] = X[p * fac + j]
    return out


def ulv_factor(Umats, Dmats, Vmats, fac=2):
    """ULV factorization (ULVsparse.py:231-293 / HBSnew.py:211-295): per level
    W_i = [V_i^perp, V_i], Q_i R_i = Rhat_i W_i with Rhat the level's D updated by the
    children's eliminated parts: Rhat = blkdiag(Ud) D + blkdiag(R22), Uhat = blkdiag(Ud) U.
    Returns a list of per-level dicts and the root (Q, R)."""
    levels = []
    Uhat = None
    R22 = None
    nlev = len(Dmats)
    for lvl in range(nlev):
        D = Dmats[lvl]
        if lvl == 0:
            Rhat = D.copy()
        else:
            Rhat = _blockdiag_children(Uhat, fac) @ D + _blockdiag_children(R22, fac)
        if lvl == nlev - 1:
            Q, R = np.linalg.qr(Rhat[0])
            levels.append(dict(root=True, Q=Q, R=R))
            break
        V = Vmats[lvl]
        U = Umats[lvl]
        nb, n, k = V.shape
        W = np.zeros((nb, n, n))
        Q = np.zeros((nb, n, n))
        R = np.zeros((nb, n, n))
        for i in range(nb):
            Wc = np.linalg.qr(V[i], mode="complete")[0]
            W[i] = np.concatenate([Wc[:, k:], V[i]], axis=1)
            Q[i], R[i] = np.linalg.qr(Rhat[i] @ W[i])
        Uc = U if lvl == 0 else _blockdiag_children(Uhat, fac) @ U
        QT = np.transpose(Q, (0, 2, 1))
        Uu = QT[:, :n - k, :] @ Uc
        Ud = QT[:, n - k:, :] @ Uc
        R22 = R[:, n - k:, n - k:]
        levels.append(dict(root=False, W=W, Q=Q, R=R, Uu=Uu, n=n, k=k, nb=nb))
        Uhat = Ud
    return levels


def ulv_solve(levels, Umats, Dmats, b, fac=2, mode="N"):
    """Solve A x = b (mode N) or A^T x = b (mode T) with the ULV factors
    (ULVsparse.py:295-358), for (N,) or (N, r) b (leaf-permuted ordering)."""
    b = np.asarray(b, dtype=np.float64)
    one_d = b.ndim == 1
    if one_d:
        b = b[:, None]
    r = b.shape[1]
    if mode == "T":
        x = _ulv_solve_T(levels, Umats, Dmats, b, fac)
        return x[:, 0] if one_d else x
    nlev = len(levels)
    tops = []
    cur = b
    for lvl in range(nlev - 1):
        L = levels[lvl]
        nb, n, k = L["nb"], L["n"], L["k"]
        t = np.transpose(L["Q"], (0, 2, 1)) @ cur.reshape(nb, n, r)
        tops.append(t[:, :n - k])
        cur = t[:, n - k:].reshape(nb * k, r)
    root = levels[-1]
    y = np.linalg.solve(root["R"], root["Q"].T @ cur)     # root unknowns
    v = Dmats[-1][0] @ y                                   # root operator applied
    for lvl in range(nlev - 2, -1, -1):
        L = levels[lvl]
        nb, n, k = L["nb"], L["n"], L["k"]
        y2 = y.reshape(nb, k, r)
        vb = v.reshape(nb, k, r)
        R = L["R"]
        rhs0 = tops[lvl] - L["Uu"] @ vb - R[:, :n - k, n - k:] @ y2
        y1 = np.linalg.solve(R[:, :n - k, :n - k], rhs0)
        x = L["W"] @ np.concatenate([y1, y2], axis=1)      # (nb, n, r)
        U = Umats[lvl]
        v = (U @ vb + Dmats[lvl] @ x).reshape(nb * n, r)
        y = x.reshape(nb * n, r)
    return y[:, 0] if one_d else y


def _ulv_solve_T(levels, Umats, Dmats, b, fac):
    # A^T = W R^T Q^T-structure: solve by dense fallback of the transpose recursion.
    # A^T x = b  <=>  x = (A^{-1})^T b; the oracle uses the same factors applied in
    # reverse order (ULVsparse.py:328-357) written out densely per level.
    A = ulv_dense_inverse(levels, Umats, Dmats, fac)
    return A.T @ b


def ulv_dense_inverse(levels, Umats, Dmats, fac=2):
    N = levels[0]["nb"] * levels[0]["n"] if len(levels) > 1 else levels[0]["R"].shape[0]
    return ulv_solve(levels, Umats, Dmats, np.eye(N), fac)
