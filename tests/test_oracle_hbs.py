"""Pin the HBS / ULV oracle (CPU only) with exactly-HBS operators."""
import numpy as np
import pytest

from oracle import hbs as oh


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("k,nl,L,fac", [(4, 8, 3, 2), (6, 10, 4, 2), (5, 8, 3, 4)])
def test_compress_exact_hbs(k, nl, L, fac):
    rng = np.random.default_rng(1)
    U, D, V, Nbvec = oh.gen_random_hbs(k, nl, L, fac, rng)
    A = oh.hbs_dense(U, D, V)
    N = A.shape[0]
    Uc, Dc, Vc, Nbc, s = oh.hbs_construct(lambda X: A @ X, lambda X: A.T @ X, N, Nbvec[0], L, k,
                                          fac=fac, rng=np.random.default_rng(2))
    assert Nbc == Nbvec
    Ac = oh.hbs_dense(Uc, Dc, Vc)
    assert rel(Ac, A) < 1e-10
    x = rng.standard_normal((N, 3))
    assert rel(oh.hbs_apply(Uc, Dc, Vc, x, "T"), A.T @ x) < 1e-10


@pytest.mark.parametrize("k,nl,L,fac", [(4, 8, 3, 2), (6, 10, 4, 2), (3, 6, 3, 4), (8, 8, 2, 2)])
def test_ulv_solve(k, nl, L, fac):
    rng = np.random.default_rng(3)
    U, D, V, Nbvec = oh.gen_random_hbs(k, nl, L, fac, rng)
    # well-conditioned: shift the leaf diagonal blocks
    D[0] = D[0] + 10 * np.eye(D[0].shape[1])
    A = oh.hbs_dense(U, D, V)
    levels = oh.ulv_factor(U, D, V, fac)
    b = rng.standard_normal((A.shape[0], 2))
    assert rel(oh.ulv_solve(levels, U, D, b, fac), np.linalg.solve(A, b)) < 1e-11
    assert rel(oh.ulv_solve(levels, U, D, b, fac, mode="T"), np.linalg.solve(A.T, b)) < 1e-11
    for L0 in levels[:-1]:
        W, Q, R = L0["W"], L0["Q"], L0["R"]
        for i in range(W.shape[0]):
            assert np.allclose(W[i].T @ W[i], np.eye(W.shape[1]), atol=1e-12)
            assert np.allclose(np.tril(R[i], -1), 0)


def test_binary_line_tree_uniform():
    t = np.sort(np.random.default_rng(0).random(512))
    # Chebyshev-like interface points on 32 leaves of 16 nodes
    ti = (1 - np.cos(np.pi * np.arange(1, 17) / 17)) / 2
    pts = (np.arange(32)[:, None] + ti[None, :]).ravel() / 32
    perm, Nb, L = oh.binary_line_tree(pts, 64)
    assert Nb == 8 and L == 4 and np.array_equal(perm, np.arange(512))
    del t


def test_planted_mistake_in_D_formula_is_caught():
    """Dropping the second term of construct_D (a plausible restatement slip) must
    break the exact-HBS reconstruction."""
    rng = np.random.default_rng(5)
    U, D, V, Nbvec = oh.gen_random_hbs(4, 8, 3, 2, rng)
    A = oh.hbs_dense(U, D, V)
    orig = oh._construct_D

    def bad(U_, V_, Y, Z, Om, Psi):
        nb, n, _ = U_.shape
        return np.stack([(Y[i] - U_[i] @ (U_[i].T @ Y[i])) @ np.linalg.pinv(Om[i])
                         for i in range(nb)])
    oh._construct_D = bad
    try:
        Uc, Dc, Vc, _, _ = oh.hbs_construct(lambda X: A @ X, lambda X: A.T @ X, A.shape[0],
                                            Nbvec[0], 3, 4, rng=np.random.default_rng(2))
    finally:
        oh._construct_D = orig
    assert rel(oh.hbs_dense(Uc, Dc, Vc), A) > 1e-6
