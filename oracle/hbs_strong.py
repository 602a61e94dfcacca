"""Oracle: strong-admissibility HBS compression from sketches, and its apply (numpy FP64,
CPU).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates matAssembly/HBS/HBStorch_strong.py (SimonDirckx/SslabLU):
  * the per-level construction of HBSMAT.constructHBS (:1033-1107): while some box has a
    far field, per box i with near field NF(i) = {i} + adj(i) (sorted, NearField
    :199-245) the near-field stack S = [Om_j ; j in NF(i)] is factored (QR of S^T,
    _factor_nf :261-287), the far-field basis U_i spans Y_i (I - Q Q^T) (leading left
    singular vectors, _range_via_projection :318-332), and the near-field error term
    E_ij = (I - U_i U_i^T) A_ij + U_i U_i^T A_ij (I - V_j V_j^T) is read from
    (I - U_i U_i^T) Y_i pinv(S) and (I - V_j V_j^T) Z_j pinv(S_Psi)
    (compute_bases_and_E_strong :687-797, _pinv_rhs :303-315);
  * the reduction to the next level: Y <- U^T (Y - E Om), Z <- V^T (Z - E^T Psi),
    Om <- V^T Om, Psi <- U^T Psi, children regrouped by `fac` (:1070-1086);
  * the dense base B = Y pinv(Om) once no box has a far field (compute_final_E :897-906);
  * the apply A = U B V^T + E per level (HBSMAT._sweep :1232-1294), mode N and T;
  * the adjacency lists it consumes (matAssembly/HBS/slabTree.py:349-545): boxes that
    share an edge ('partial') or touch ('full'), here by brute force over all pairs of
    boxes of a level, from the box bounds.

basis='svd' is the reference's choice of far-field basis; basis='pqr' takes the first rk
columns of a column-pivoted QR of the Gram matrix M M^T of the same projected sketch M
instead (the device engine's choice); both span the far field exactly when it has rank
<= rk.

Parity pins (tests/test_oracle_hbs_strong.py): compression of operators that are exactly
strong-admissible at rank rk (nested bases, random dense near field; the construction of
HBSnew_strong.py:528-566 generalised to any tree adjacency) reproduces them to ~1e-12 in
modes N and T on line, rectangle and quad trees; the leaf-level E tiles equal the closed
form A_ij - U_i U_i^T A_ij V_j V_j^T from the dense matrix; far-field blocks satisfy
(I - U_i U_i^T) A_ij = 0; planted mistakes (dropping E's second term, skipping the
near-field projection) break the exact recovery.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla


# ---------------------------------------------------------------------------
# adjacency by brute force

def touching(bp, bq, edge_only, alpha=1e-3):
    """Boxes (lo, hi per axis, flattened) touch: the gap on every axis is <= alpha times
    the smaller width.  edge_only: additionally they overlap with positive length along
    all but one axis (share an edge, not only a corner)."""
    nax = len(bp) // 2
    close = 0
    for a in range(nax):
        lo_p, hi_p, lo_q, hi_q = bp[2 * a], bp[2 * a + 1], bq[2 * a], bq[2 * a + 1]
        h = min(hi_p - lo_p, hi_q - lo_q)
        gap = max(lo_p - hi_q, lo_q - hi_p)
        if gap > alpha * h:
            return False
        if gap > -alpha * h:       # only meets the other box at a point on this axis
            close += 1
    return close <= 1 if edge_only else True


def brute_adjacency(bounds_by_level, edge_only):
    """bounds_by_level[l]: list of box bounds in rank order.  Returns per level the list
    of neighbour sets (self excluded)."""
    out = []
    for bl in bounds_by_level:
        n = len(bl)
        out.append([{j for j in range(n) if j != i and touching(bl[i], bl[j], edge_only)}
                    for i in range(n)])
    return out


# ---------------------------------------------------------------------------
# exactly strong-admissible operators

def exact_strong_operator(adj, b, r, fac, seed=0):
    """Dense (Nb b x Nb b) operator that the strong format holds exactly at rank r.

    adj[l]: neighbour sets of the boxes of level l (level-local ranks, children of rank P
    at level l are ranks fac P .. fac P + fac - 1 at level l + 1), leaf level last.
    Nested bases: leaf U_i (b x r) orthonormal, a parent's basis is the block diagonal of
    its children's times a (fac r x r) transfer matrix.  Blocks (P, Q) of level l that
    are not neighbours but whose parents are (or are the same box) get hU_P S hV_Q^T;
    leaf neighbours (and the diagonal) get dense random blocks."""
    rng = np.random.default_rng(seed)
    L = len(adj) - 1
    Nb = len(adj[L])
    N = Nb * b
    orth = lambda m, k: np.linalg.qr(rng.standard_normal((m, k)))[0]  # noqa: E731
    hU, hV = {}, {}
    for i in range(Nb):
        hU[(L, i)], hV[(L, i)] = orth(b, r), orth(b, r)
    for lvl in range(L - 1, 0, -1):
        for P in range(len(adj[lvl])):
            kids = range(fac * P, fac * P + fac)
            for h, store in ((hU, hU), (hV, hV)):
                B = sla.block_diag(*[h[(lvl + 1, c)] for c in kids])
                store[(lvl, P)] = B @ orth(fac * r, r)
    span = lambda lvl, P: (P * b * fac ** (L - lvl), (P + 1) * b * fac ** (L - lvl))  # noqa
    A = np.zeros((N, N))
    for lvl in range(1, L + 1):
        n = len(adj[lvl])
        for P in range(n):
            for Q in range(n):
                near = P == Q or Q in adj[lvl][P]
                pp, pq = P // fac, Q // fac
                parent_near = pp == pq or pq in adj[lvl - 1][pp]
                if not near and parent_near:
                    r0, r1 = span(lvl, P)
                    c0, c1 = span(lvl, Q)
                    A[r0:r1, c0:c1] = hU[(lvl, P)] @ rng.standard_normal((r, r)) @ hV[(lvl, Q)].T
    for i in range(Nb):
        for j in [i] + sorted(adj[L][i]):
            A[i * b:(i + 1) * b, j * b:(j + 1) * b] = rng.standard_normal((b, b))
    return A


# ---------------------------------------------------------------------------
# construction (HBStorch_strong.py:1033-1107)

def _stack_pinv(S):
    """(Q, Rinv^T) with S^T = Q R (reduced), so pinv(S) = Q R^{-T} (_pinv_rhs :303-315);
    needs full row rank (k b <= s, Gaussian test matrices)."""
    Q, R = np.linalg.qr(S.T)
    return Q, np.linalg.inv(R).T


def _far_basis(M, rk, basis):
    if basis == "svd":
        return np.linalg.svd(M, full_matrices=False)[0][:, :rk]
    Qm, _, _ = sla.qr(M @ M.T, pivoting=True)
    return Qm[:, :rk]


def has_far(adj_l):
    n = len(adj_l)
    return any(len(adj_l[i] | {i}) < n for i in range(n))


def construct(Om, Psi, Y, Z, adj, rk, fac=2, basis="svd", _plant=None):
    """Om, Psi (N x s) test matrices and Y = A Om, Z = A^T Psi in leaf order; adj as in
    exact_strong_operator.  Returns dict(levels=[(U, V, E, nf)], B, fac, b0)."""
    L = len(adj) - 1
    nb = len(adj[L])
    s = Om.shape[1]
    blk = lambda M, n: [M[i * (M.shape[0] // n):(i + 1) * (M.shape[0] // n)] for i in range(n)]  # noqa
    Ob, Pb, Yb, Zb = (blk(M, nb) for M in (Om, Psi, Y, Z))
    levels = []
    lvl = L
    while has_far(adj[lvl]):
        nb = len(adj[lvl])
        b = Ob[0].shape[0]
        rkm = min(rk, b)
        nf = [sorted(adj[lvl][i] | {i}) for i in range(nb)]
        U, V, X1, X2 = [None] * nb, [None] * nb, [None] * nb, [None] * nb
        for side in ("U", "V"):
            S_b, K_b, B_out, X_out = (Ob, Yb, U, X1) if side == "U" else (Pb, Zb, V, X2)
            for i in range(nb):
                S = np.vstack([S_b[j] for j in nf[i]])
                Q, RinvT = _stack_pinv(S)
                Ki = K_b[i]
                M = Ki if _plant == "no_projection" else Ki - (Ki @ Q) @ Q.T
                Bi = _far_basis(M, rkm, basis)
                B_out[i] = Bi
                X_out[i] = ((Ki - Bi @ (Bi.T @ Ki)) @ Q) @ RinvT   # (b, k b)
        E = {}
        for i in range(nb):
            for a, j in enumerate(nf[i]):
                E[(i, j)] = X1[i][:, a * b:(a + 1) * b].copy()
        if _plant != "drop_term2":
            for j in range(nb):
                for a, i in enumerate(nf[j]):
                    T = X2[j][:, a * b:(a + 1) * b].T                 # A_ij (I - V_j V_j^T)
                    E[(i, j)] += U[i] @ (U[i].T @ T)
        levels.append((U, V, E, nf))
        # reduction to the parent level
        Yr = [U[i].T @ (Yb[i] - sum(E[(i, j)] @ Ob[j] for j in nf[i])) for i in range(nb)]
        Zr = [V[j].T @ (Zb[j] - sum(E[(i, j)].T @ Pb[i] for i in nf[j])) for j in range(nb)]
        Or = [V[i].T @ Ob[i] for i in range(nb)]
        Pr = [U[i].T @ Pb[i] for i in range(nb)]
        grp = lambda T: [np.vstack(T[fac * p:fac * p + fac]) for p in range(nb // fac)]  # noqa
        Yb, Zb, Ob, Pb = grp(Yr), grp(Zr), grp(Or), grp(Pr)
        lvl -= 1
    Yf, Of = np.vstack(Yb), np.vstack(Ob)
    B = Yf @ np.linalg.pinv(Of)
    return dict(levels=levels, B=B, fac=fac, nb_base=len(Yb), s=s)


def apply(H, x, mode="N"):
    """y = A x (mode N) or A^T x (mode T) in leaf order (HBSMAT._sweep :1232-1294)."""
    x = np.asarray(x, dtype=np.float64)
    one = x.ndim == 1
    X = x[:, None] if one else x
    fac = H["fac"]
    down = []
    g = X
    for U, V, E, nf in H["levels"]:
        nb = len(U)
        gb = np.split(g, nb)
        down.append(gb)
        D = V if mode == "N" else U
        g = np.vstack([D[i].T @ gb[i] for i in range(nb)])
    w = (H["B"] if mode == "N" else H["B"].T) @ g
    for (U, V, E, nf), gb in zip(reversed(H["levels"]), reversed(down)):
        nb = len(U)
        Up = U if mode == "N" else V
        wb = np.split(w, nb)
        out = []
        for i in range(nb):
            near = np.zeros_like(gb[i])
            for j in nf[i]:
                near += (E[(i, j)] if mode == "N" else E[(j, i)].T) @ gb[j]
            out.append(Up[i] @ wb[i] + near)
        w = np.vstack(out)
    return w[:, 0] if one else w


def dense(H, N):
    return apply(H, np.eye(N))


def sample_count(adj, b, rk, p=20, fac=2):
    """Samples for a given adjacency: max near-field size times max(fac rk, b) + rk + p
    (the count test/rkprop/thinSlab3D.py:400-408 uses with its kmax, for fac = 2)."""
    kmax = max(max((len(a) + 1 for a in al), default=1) for al in adj)
    return kmax * max(fac * rk, b) + rk + p
