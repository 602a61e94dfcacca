"""Oracle: block-tridiagonal S-system solvers (numpy/scipy FP64, CPU).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates
  * Thomas with general diagonal T and Sherman-Morrison-Woodbury corner correction
    -- direct_solve/omsdirectsolve.py:48-175 (BlockTridiagonalSolver)
  * red-black cyclic reduction -- direct_solve/omsdirectsolve.py:177-333 (RedBlackSolver)
  * identity-diagonal Thomas / cyclic SMW -- multislab/omsdirectsolve.py:17-160
  * dense assembly of the block system -- direct_solve/solver_test_helpers.py:13-43
  * the random diagonally-dominant test systems -- direct_solve/test_RB.py:14-45

Parity pins (tests/test_oracle_blocktri.py):
  * every solver against ``numpy.linalg.solve`` of the assembled dense matrix (the
    closed form used by the reference's own test, direct_solve/test_RB.py:51-60),
    including the nine (nSlabs, m) cases of test_RB.py:70-81 at tolerance 1e-8 and
    cyclic systems assembled as in solver_test_helpers.py:22-31;
  * red-black with nb = 1 reduces to a single LU solve; Thomas and red-black agree.
Differences from the reference kept on purpose: red-black accepts (N, r) right-hand
sides (columns are independent) and any block count (identity padding to the next
power of two, which is exact); the reference's dense red-black guard builds but never
raises its ValueError (omsdirectsolve.py:201-202).
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lu_factor, lu_solve


# ---------------------------------------------------------------------------
# generators / assembly

def make_block_tridiagonal(nSlabs: int, m: int, seed: int = 42):
    """Random diagonally-dominant block-tridiagonal system, drawing from
    ``np.random.default_rng(seed)`` in the order of direct_solve/test_RB.py:24-44
    (per block row: the diagonal block, then the sub- and super-diagonal blocks
    when they exist, then one right-hand side of length nSlabs*m).

    Returns (S_rk_list, T, dense_matrix, rhs) with S_rk_list[i] = (A_i, C_i)."""
    rng = np.random.default_rng(seed)
    N = nSlabs * m
    dense = np.zeros((N, N))
    T, S = [], []
    for i in range(nSlabs):
        diag = rng.standard_normal((m, m)) + nSlabs * m * np.eye(m)
        T.append(diag)
        dense[i * m:(i + 1) * m, i * m:(i + 1) * m] = diag
        sub = 0.1 * rng.standard_normal((m, m)) if i > 0 else np.zeros((m, m))
        sup = 0.1 * rng.standard_normal((m, m)) if i < nSlabs - 1 else np.zeros((m, m))
        if i > 0:
            dense[i * m:(i + 1) * m, (i - 1) * m:i * m] = sub
        if i < nSlabs - 1:
            dense[i * m:(i + 1) * m, (i + 1) * m:(i + 2) * m] = sup
        S.append((sub, sup))
    rhs = rng.standard_normal(N)
    return S, T, dense, rhs


def assemble_dense(S_rk_list, T, cyclic=False):
    """Dense matrix of the block system (solver_test_helpers.py:22-31):
    row block i holds T[i] on the diagonal, S_rk_list[i][0] at block column i-1 and
    S_rk_list[i][-1] at block column i+1 (wrapping around when cyclic)."""
    nb = len(T)
    m = T[0].shape[0]
    A = np.zeros((nb * m, nb * m))
    for i in range(nb):
        A[i * m:(i + 1) * m, i * m:(i + 1) * m] = T[i]
        if i != 0 or cyclic:
            j = (i - 1) % nb
            A[i * m:(i + 1) * m, j * m:(j + 1) * m] += S_rk_list[i][0]
        if i != nb - 1 or cyclic:
            j = (i + 1) % nb
            A[i * m:(i + 1) * m, j * m:(j + 1) * m] += S_rk_list[i][-1]
    return A


# ---------------------------------------------------------------------------
# Thomas (direct_solve/omsdirectsolve.py:48-175)

def thomas_factor(S_rk_list, T):
    """Forward elimination factors (omsdirectsolve.py:77-100):
    A_i = S_{i,i-1} B_{i-1}^{-1} (explicit inverse through lu_solve(B, I)),
    B_i = lu(T_i - A_i C_{i-1}), C_i = S_rk_list[i][-1].  Returns (A, B, C)."""
    m = S_rk_list[0][0].shape[0]
    n = len(S_rk_list) - 1
    eye = np.eye(m)
    A = []
    B = [lu_factor(T[0])]
    C = [S_rk_list[k][-1] for k in range(n)]
    for i in range(1, n + 1):
        Ai = S_rk_list[i][0] @ lu_solve(B[-1], eye)
        B.append(lu_factor(T[i] - Ai @ C[i - 1]))
        A.append(Ai)
    return A, B, C


def thomas_solve(A, B, C, rhs, indices=None, m=None):
    """Forward sweep d_i -= A_i d_{i-1}, then back substitution
    x_i = B_i \\ (d_i - C_i x_{i+1})  (omsdirectsolve.py:102-141)."""
    n = len(A)
    if m is None:
        m = B[0][0].shape[0]
    if indices is None:
        indices = [range(k * m, (k + 1) * m) for k in range(n + 1)]
    d = np.array(rhs, dtype=np.float64, copy=True)
    for i in range(1, n + 1):
        d[indices[i]] = d[indices[i]] - A[i - 1] @ d[indices[i - 1]]
    x = np.zeros(d.shape)
    x[indices[n]] = lu_solve(B[n], d[indices[n]])
    for i in range(n - 1, -1, -1):
        x[indices[i]] = lu_solve(B[i], d[indices[i]] - C[i] @ x[indices[i + 1]])
    return x


def thomas_cyclic_smw(S_rk_list, factors, indices=None):
    """Woodbury correction for the corner blocks E = S_rk_list[0][0] (row 0,
    column nb-1) and F = S_rk_list[-1][-1] (row nb-1, column 0):
    smw = Z (I + V^T Z)^{-1} V^T with Z = T^{-1} U (omsdirectsolve.py:148-166)."""
    E = S_rk_list[0][0]
    F = S_rk_list[-1][-1]
    m = E.shape[0]
    n = len(S_rk_list)
    U = np.zeros((n * m, 2 * m))
    V = np.zeros((n * m, 2 * m))
    U[0:m, 0:m] = E
    U[(n - 1) * m:n * m, m:2 * m] = F
    V[(n - 1) * m:n * m, 0:m] = np.eye(m)
    V[0:m, m:2 * m] = np.eye(m)
    Z = thomas_solve(*factors, U, indices, m)
    return Z @ np.linalg.solve(np.eye(2 * m) + V.T @ Z, V.T)


def thomas_direct(S_rk_list, T, rhs, cyclic=False):
    """Factor + solve (+ cyclic correction, omsdirectsolve.py:169-175)."""
    fac = thomas_factor(S_rk_list, T)
    x = thomas_solve(*fac, rhs)
    if cyclic:
        x = x - thomas_cyclic_smw(S_rk_list, fac) @ x
    return x


def identity_thomas_factor(S_rk_list):
    """Identity-diagonal Thomas of multislab/omsdirectsolve.py:17-68:
    B_0 = I, A_1 = S_{1,0}, A_i = S_{i,i-1} B_{i-1}^{-1}, B_i = lu(I - A_i C_{i-1})."""
    m = S_rk_list[0][0].shape[0]
    n = len(S_rk_list) - 1
    eye = np.eye(m)
    A, B = [], [eye]
    C = [S_rk_list[k][-1] for k in range(n)]
    for i in range(1, n + 1):
        if i == 1:
            Ai = S_rk_list[i][0] @ B[-1]
        else:
            Ai = S_rk_list[i][0] @ lu_solve(B[-1], eye)
        B.append(lu_factor(eye - Ai @ C[i - 1]))
        A.append(Ai)
    return A, B, C


def identity_thomas_solve(factors, rhs, indices):
    """multislab/omsdirectsolve.py:70-99 (B_0 = I needs no solve)."""
    A, B, C = factors
    n = len(A)
    d = np.array(rhs, dtype=np.float64, copy=True)
    for i in range(1, n + 1):
        d[indices[i]] = d[indices[i]] - A[i - 1] @ d[indices[i - 1]]
    x = np.zeros(d.shape)
    x[indices[n]] = lu_solve(B[n], d[indices[n]])
    for i in range(n - 1, 0, -1):
        x[indices[i]] = lu_solve(B[i], d[indices[i]] - C[i] @ x[indices[i + 1]])
    x[indices[0]] = d[indices[0]] - C[0] @ x[indices[1]]
    return x


# ---------------------------------------------------------------------------
# Red-black (direct_solve/omsdirectsolve.py:177-333)

def _lu_to_dense(f):
    """Rebuild T from its lu_factor pair (reconstruct_from_lu_factor,
    omsdirectsolve.py:31-46): apply the recorded row swaps to L U in reverse."""
    lu, piv = f
    n = lu.shape[0]
    M = (np.tril(lu, -1) + np.eye(n)) @ np.triu(lu)
    for i in range(n - 1, -1, -1):
        p = piv[i]
        if p != i:
            M[[i, p]] = M[[p, i]]
    return M


def redblack_level(m, nSlabs, level, cyclic=False):
    """One reduction level (omsdirectsolve.py:283-333): for even i,
    B_i = T_i - SiM_i T_{i-1}^{-1} SiP_{i-1} - SiP_i T_{i+1}^{-1} SiM_{i+1},
    A_i = -SiM_i T_{i-1}^{-1} SiM_{i-1},  C_i = -SiP_i T_{i+1}^{-1} SiP_{i+1};
    at the coarsest cyclic level B_0 += A_0 + C_0."""
    SiM, _, T_lu, SiP = level
    eye = np.eye(m)
    Bn, An, Cn = [], [], []
    for i in range(0, nSlabs, 2):
        Bi = _lu_to_dense(T_lu[i])
        if i > 0 or cyclic:
            Bi = Bi - SiM[i] @ lu_solve(T_lu[i - 1], SiP[(i - 1) % nSlabs])
        if i < nSlabs - 1 or cyclic:
            Bi = Bi - SiP[i] @ lu_solve(T_lu[i + 1], SiM[(i + 1) % nSlabs])
        Bn.append(Bi)
        An.append(-SiM[i] @ lu_solve(T_lu[i - 1], SiM[(i - 1) % nSlabs]))
        Cn.append(-SiP[i] @ lu_solve(T_lu[i + 1], SiP[(i + 1) % nSlabs]))
    if len(Bn) == 1 and cyclic:
        Bn[0] = Bn[0] + An[0] + Cn[0]
    return (An, Bn, [lu_factor(b) for b in Bn], Cn)


def redblack_factor(S_rk_list, T, cyclic=False):
    """Red-black factorization (omsdirectsolve.py:180-219).  nSlabs must be a power of
    two; non-cyclic systems ignore S_rk_list[0][0] and S_rk_list[-1][-1]."""
    m = S_rk_list[0][0].shape[0]
    nSlabs = len(S_rk_list)
    if nSlabs & (nSlabs - 1):
        raise ValueError("red-black oracle needs a power-of-two block count; pad first")
    SiM = [s[0] for s in S_rk_list]
    SiP = [s[-1] for s in S_rk_list]
    if not cyclic:
        SiM[0] = np.zeros((m, m))
        SiP[-1] = np.zeros((m, m))
    RB = [(SiM, list(T), [lu_factor(t) for t in T], SiP)]
    n = nSlabs
    while n > 1:
        RB.append(redblack_level(m, n, RB[-1], cyclic))
        n //= 2
    return RB


def redblack_solve(RB, rhs, cyclic=False):
    """Forward reduction then back substitution over the levels
    (omsdirectsolve.py:221-279), for (N,) or (N, r) right-hand sides."""
    m = RB[0][0][0].shape[0]
    v = np.array(rhs, dtype=np.float64, copy=True)
    one_d = v.ndim == 1
    if one_d:
        v = v[:, None]
    r = v.shape[1]
    vs = [v]
    for lvl in range(len(RB) - 1):
        SiM, _, T_lu, SiP = RB[lvl]
        n = len(SiM)
        nr = n // 2
        prev_v = vs[-1]
        new_v = np.zeros((m * nr, r))
        for j in range(nr):
            i = 2 * j
            c = prev_v[i * m:(i + 1) * m].copy()
            if cyclic or i > 0:
                p = (i - 1) % n
                c -= SiM[i] @ lu_solve(T_lu[p], prev_v[p * m:(p + 1) * m])
            if cyclic or i < n - 1:
                q = (i + 1) % n
                c -= SiP[i] @ lu_solve(T_lu[q], prev_v[q * m:(q + 1) * m])
            new_v[j * m:(j + 1) * m] = c
        vs.append(new_v)
    vs[-1] = lu_solve(RB[-1][2][0], vs[-1])
    for lvl in range(len(RB) - 1, 0, -1):
        SiM, _, T_lu, SiP = RB[lvl - 1]
        nr = len(SiM) // 2
        cur = vs[lvl - 1]
        coarse = vs[lvl]
        for j in range(nr):
            i = 2 * j
            cur[i * m:(i + 1) * m] = coarse[j * m:(j + 1) * m]
        for j in range(nr):
            i = 2 * j
            nxt = (j + 1) % nr
            c = SiM[i + 1] @ coarse[j * m:(j + 1) * m]
            if cyclic or j + 1 < nr:
                c = c + SiP[i + 1] @ coarse[nxt * m:(nxt + 1) * m]
            cur[(i + 1) * m:(i + 2) * m] = lu_solve(T_lu[i + 1], cur[(i + 1) * m:(i + 2) * m] - c)
    x = vs[0]
    return x[:, 0] if one_d else x


def pad_pow2(S_rk_list, T):
    """Pad a non-cyclic system to the next power-of-two block count with identity
    diagonal blocks and zero couplings (exact: the padded unknowns are zero)."""
    nb = len(T)
    m = T[0].shape[0]
    n2 = 1
    while n2 < nb:
        n2 *= 2
    S = [(s[0], s[-1]) for s in S_rk_list]
    S[-1] = (S[-1][0], np.zeros((m, m)))
    TT = list(T)
    for _ in range(nb, n2):
        S.append((np.zeros((m, m)), np.zeros((m, m))))
        TT.append(np.eye(m))
    return S, TT, n2


def redblack_direct(S_rk_list, T, rhs, cyclic=False):
    """Factor + solve with padding for non-power-of-two non-cyclic systems."""
    nb = len(T)
    m = T[0].shape[0]
    if nb & (nb - 1):
        if cyclic:
            raise ValueError("cyclic red-black oracle needs a power-of-two block count")
        S, TT, n2 = pad_pow2(S_rk_list, T)
        v = np.array(rhs, dtype=np.float64)
        pad_shape = ((n2 - nb) * m,) + v.shape[1:]
        vv = np.concatenate([v, np.zeros(pad_shape)], axis=0)
        x = redblack_solve(redblack_factor(S, TT), vv)
        return x[:nb * m]
    return redblack_solve(redblack_factor(S_rk_list, T, cyclic), rhs, cyclic)
