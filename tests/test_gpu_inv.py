"""GPU: batched explicit inverses (sslab_dinv_batched), the red-black factor's T^{-1}.

Block elimination down to 64 x 64 Gauss-Jordan blocks, with a sampled residual check that
falls back to the pivoted LU.  Pins: the closed-form inverse of the (-1, 2, -1) matrix,
numpy's inverse on well-conditioned batches (padded leading dimensions and strides, sizes
across the 64 / 96 / 128 recursion boundaries), matrices whose leading block is singular
(must take the LU path and still be exact), and exactly singular matrices (LU info).
Tolerance: relative error <= 1e-12 (FP64) on matrices of condition number <= 1e4.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch, gpu


def ginv(env, A, lda=None, ldx=None, tol=0.0):
    """Invert the (batch, n, n) array A through the C ABI; returns (X, info, used_lu)."""
    torch, dev = env
    from sslablu_amd._lib import call, ptr, sizes2, stream_handle
    b, n, _ = A.shape
    lda = lda or n
    ldx = ldx or n
    Ap = np.zeros((b, n, lda))
    Ap[:, :, :n] = A
    Ad = torch.tensor(Ap, device=dev)
    Xd = torch.full((b, n, ldx), 7.0, dtype=torch.float64, device=dev)
    w, iw = sizes2("sslab_inv_sizes", n, b)
    work = torch.empty(max(w, 1), dtype=torch.float64, device=dev)
    iwork = torch.empty(max(iw, 1), dtype=torch.int32, device=dev)
    info = torch.full((b,), -5, dtype=torch.int32, device=dev)
    used = ctypes.c_int(-1)
    call("sslab_dinv_batched", n, ptr(Ad), lda, n * lda, ptr(Xd), ldx, n * ldx, b, ptr(work),
         ptr(iwork), ptr(info), float(tol), ctypes.addressof(used), stream_handle())
    torch.cuda.synchronize()
    X = Xd.cpu().numpy()
    assert np.all(X[:, :, n:] == 7.0), "columns past n of X were written"
    return X[:, :, :n], info.cpu().numpy(), used.value


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("n", [1, 7, 64, 65, 96, 97, 130, 200, 511])
def test_second_difference_closed_form(env, n):
    """(tridiag(-1, 2, -1))^{-1}_{ij} = min(i, j) (n + 1 - max(i, j)) / (n + 1), 1-based."""
    A = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    i = np.arange(1, n + 1)
    Xc = np.minimum.outer(i, i) * (n + 1 - np.maximum.outer(i, i)) / (n + 1)
    X, info, used = ginv(env, A[None])
    assert used == 0 and np.all(info == 0)
    assert rel(X[0], Xc) <= 1e-12


@pytest.mark.parametrize("n,batch,lda,ldx", [(64, 5, 64, 64), (100, 3, 103, 101),
                                             (256, 4, 256, 260), (512, 2, 512, 512),
                                             (333, 2, 340, 333)])
def test_well_conditioned_batches_vs_numpy(env, n, batch, lda, ldx):
    rng = np.random.default_rng(n + batch)
    A = np.eye(n) + 0.4 * rng.standard_normal((batch, n, n)) / np.sqrt(n)
    X, info, used = ginv(env, A, lda, ldx)
    assert used == 0 and np.all(info == 0)
    for b in range(batch):
        assert np.linalg.cond(A[b]) < 1e4
        assert rel(X[b], np.linalg.inv(A[b])) <= 1e-12


def test_singular_leading_block_takes_the_lu_path(env):
    """[[0, B], [C, 0]]: the 64 x 64 leading block is zero, so block elimination breaks
    down; the sampled residual must reject it and the pivoted LU must give the inverse.
    The well-conditioned matrix in the same batch is inverted by the same LU pass."""
    rng = np.random.default_rng(3)
    n, h = 128, 64
    B = np.eye(h) + 0.3 * rng.standard_normal((h, h)) / np.sqrt(h)
    C = np.eye(h) + 0.3 * rng.standard_normal((h, h)) / np.sqrt(h)
    A0 = np.block([[np.zeros((h, h)), B], [C, np.zeros((h, h))]])
    A1 = np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    X, info, used = ginv(env, np.stack([A1, A0]))
    assert used == 1 and np.all(info == 0)
    assert rel(X[0], np.linalg.inv(A1)) <= 1e-12
    assert rel(X[1], np.linalg.inv(A0)) <= 1e-12


def test_exactly_singular_reports_lu_info(env):
    rng = np.random.default_rng(4)
    n = 90
    A = np.eye(n) + 0.3 * rng.standard_normal((2, n, n)) / np.sqrt(n)
    A[1, 40, :] = 0.0  # a zero row: U[j][j] = 0 for some j
    _, info, used = ginv(env, A)
    assert used == 1
    assert info[0] == 0 and info[1] > 0


def test_tolerance_forces_the_lu_path(env):
    """A tolerance no residual meets sends even a well-conditioned batch to the LU."""
    rng = np.random.default_rng(5)
    n = 150
    A = np.eye(n) + 0.3 * rng.standard_normal((2, n, n)) / np.sqrt(n)
    X, info, used = ginv(env, A, tol=1e-300)
    assert used == 1 and np.all(info == 0)
    for b in range(2):
        assert rel(X[b], np.linalg.inv(A[b])) <= 1e-12


@pytest.mark.parametrize("n", [5, 50, 64])
def test_row_pivoting_inside_a_block(env, n):
    """A permuted, well-conditioned matrix with a zero diagonal: every Gauss-Jordan step
    inside the 64 x 64 base block must pivot away from the diagonal row."""
    rng = np.random.default_rng(n)
    P = np.eye(n)[rng.permutation(n)]
    while np.any(np.diag(P) != 0):  # a derangement: no row may pivot on its own diagonal
        P = np.eye(n)[rng.permutation(n)]
    A = (P + 0.05 * rng.standard_normal((n, n)) / np.sqrt(n)) * (1 - np.eye(n))
    X, info, used = ginv(env, A[None])
    assert used == 0 and np.all(info == 0)
    assert rel(X[0], np.linalg.inv(A)) <= 1e-12


def test_scaled_matrices_pass_the_backward_error_check(env):
    """The check is scale-free: matrices of norm 1e8 and 1e-6 (a spectral leaf operator
    scales with p^4 / h^2) are inverted by block elimination, without the LU fallback."""
    rng = np.random.default_rng(6)
    n = 200
    B = np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    A = np.stack([1e8 * B, 1e-6 * B])
    X, info, used = ginv(env, A)
    assert used == 0 and np.all(info == 0)
    for b in range(2):
        assert rel(X[b], np.linalg.inv(A[b])) <= 1e-12
