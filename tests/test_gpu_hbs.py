"""GPU parity of the HBS engine: batched QR, compression of exactly-HBS operators,
telescoping apply and ULV solves against the oracle (oracle/hbs.py).

Tolerances: exactly-HBS reconstruction 1e-9 relative (randomized compression in FP64),
ULV solve 1e-10 relative to numpy.linalg.solve of the dense HBS matrix.
"""
import numpy as np
import pytest

from oracle import hbs as oh

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else a
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


@pytest.mark.parametrize("m,n,pivot", [(7, 5, False), (40, 40, False), (293, 72, False),
                                       (72, 72, True), (50, 30, True), (30, 50, True),
                                       (130, 72, False), (72, 58, True), (226, 40, False),
                                       (300, 200, True), (160, 100, True), (192, 90, True),
                                       (150, 150, True), (180, 120, False)])
def test_batched_qr(T, m, n, pivot):
    from sslablu_amd.hbs import bqr
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((3, m, n))
    if pivot:  # rank-deficient columns to exercise the pivoting
        A[:, :, n // 2:] = A[:, :, : n - n // 2] @ rng.standard_normal((n - n // 2, n - n // 2))
    Q, R, perm = bqr(T.tensor(A, device="cuda"), pivot=pivot)
    Q, R = Q.cpu().numpy(), R.cpu().numpy()
    for b in range(3):
        Ab = A[b][:, perm[b].cpu().numpy()] if pivot else A[b]
        assert np.allclose(Q[b].T @ Q[b], np.eye(m), atol=1e-12)
        assert np.allclose(np.tril(R[b], -1), 0.0)
        assert np.linalg.norm(Q[b] @ R[b] - Ab) / np.linalg.norm(Ab) < 1e-13
        if pivot:
            d = np.abs(np.diag(R[b]))
            assert np.all(d[:-1] + 1e-9 * d[0] >= d[1:])  # non-increasing |R_jj|


@pytest.mark.parametrize("m,n", [(48, 40), (160, 100), (192, 90)])
def test_pivoted_qr_ties_take_first_index(T, m, n):
    """Duplicated columns have equal norms at every step until one of a pair is chosen
    (then the other's residual vanishes): the pivot takes the first index on ties, so
    the first n/2 pivots are exactly the first half of the columns."""
    from sslablu_amd.hbs import bqr
    rng = np.random.default_rng(m)
    h = n // 2
    B = rng.standard_normal((2, m, h)) * (1.0 + np.arange(h) % 3)  # some equal scales too
    A = np.concatenate([B, B], axis=2)
    Q, R, perm = bqr(T.tensor(A, device="cuda"), pivot=True)
    perm = perm.cpu().numpy()
    for b in range(2):
        assert sorted(perm[b, :h].tolist()) == list(range(h)), perm[b]
        assert perm[b, 0] == int(np.argmax(np.linalg.norm(B[b], axis=0)))
        Ab = A[b][:, perm[b]]
        assert np.linalg.norm(Q.cpu().numpy()[b] @ R.cpu().numpy()[b] - Ab) / np.linalg.norm(Ab) < 1e-13


def _device_hbs(T, U, D, V, fac):
    from sslablu_amd.hbs import DeviceHBS
    t = lambda Xs: [T.tensor(X, device="cuda") for X in Xs]
    return DeviceHBS(t(U), t(D), t(V), fac)


@pytest.mark.parametrize("k,nl,L,fac", [(4, 8, 3, 2), (6, 10, 4, 2), (5, 8, 3, 4),
                                        (16, 16, 4, 2)])
def test_apply_and_compress(T, k, nl, L, fac):
    from sslablu_amd.hbs import DeviceHBS
    rng = np.random.default_rng(7)
    U, D, V, Nbvec = oh.gen_random_hbs(k, nl, L, fac, rng)
    A = oh.hbs_dense(U, D, V)
    H = _device_hbs(T, U, D, V, fac)
    x = rng.standard_normal((A.shape[0], 3))
    xt = T.tensor(x, device="cuda")
    assert rel(H.apply(xt), A @ x) < 1e-13
    assert rel(H.apply(xt, "T"), A.T @ x) < 1e-13
    # compression from sketches of the dense operator (same test matrices as the oracle)
    N = A.shape[0]
    s = 6 * k + 4 * nl + 5 if fac == 4 else 4 * k + 2 * nl + 5
    Om = rng.standard_normal((N, s))
    Ps = rng.standard_normal((N, s))
    dev = lambda X: T.tensor(X, device="cuda")
    Hc = DeviceHBS.from_sketches(dev(Om), dev(Ps), dev(A @ Om), dev(A.T @ Ps), Nbvec[0], L, k,
                                 fac)
    assert rel(Hc.dense(), A) < 1e-9


@pytest.mark.parametrize("k,nl,L,fac", [(4, 8, 3, 2), (6, 10, 4, 2), (3, 6, 3, 4),
                                        (8, 8, 2, 2), (16, 32, 4, 2)])
def test_ulv_solve(T, k, nl, L, fac):
    rng = np.random.default_rng(11)
    U, D, V, Nbvec = oh.gen_random_hbs(k, nl, L, fac, rng)
    D[0] = D[0] + 10 * np.eye(D[0].shape[1])
    A = oh.hbs_dense(U, D, V)
    H = _device_hbs(T, U, D, V, fac)
    b = rng.standard_normal((A.shape[0], 4))
    bt = T.tensor(b, device="cuda")
    assert rel(H.solve(bt), np.linalg.solve(A, b)) < 1e-10
    assert rel(H.solve(bt, "T"), np.linalg.solve(A.T, b)) < 1e-10
    assert rel(H.solve(bt[:, 0]), np.linalg.solve(A, b[:, 0])) < 1e-10
    # both modes from one stacked factorization (ulv_factor_pair) solve the same systems
    H2 = _device_hbs(T, U, D, V, fac)
    H2.ulv_factor_pair()
    assert set(H2._ulv) == {"N", "T"}
    assert rel(H2.solve(bt), np.linalg.solve(A, b)) < 1e-10
    assert rel(H2.solve(bt, "T"), np.linalg.solve(A.T, b)) < 1e-10
    assert rel(H2.solve(bt, "T"), H.solve(bt, "T").cpu().numpy()) < 1e-12


@pytest.mark.parametrize("permuted", [False, True])
def test_planned_solve_and_apply_replay(T, permuted, monkeypatch):
    """DeviceHBS solve / apply through launch plans: after the first call of a (mode,
    width) the GEMM launches are replayed on plan buffers; every later call with new data
    matches the dense operator, results are fresh tensors (an earlier result is not
    overwritten), and the unplanned path gives the same numbers."""
    from sslablu_amd import hbs
    rng = np.random.default_rng(21)
    k, nl, L, fac = 6, 10, 4, 2
    U, D, V, Nbvec = oh.gen_random_hbs(k, nl, L, fac, rng)
    D[0] = D[0] + 10 * np.eye(D[0].shape[1])
    A = oh.hbs_dense(U, D, V)
    N = A.shape[0]
    t = lambda Xs: [T.tensor(X, device="cuda") for X in Xs]
    perm = rng.permutation(N) if permuted else None
    H = hbs.DeviceHBS(t(U), t(D), t(V), fac, perm)
    # with a leaf permutation p the device operator acts on x as A[inv, :][:, inv]
    if permuted:
        inv = np.argsort(perm)
        A = A[np.ix_(inv, inv)]
    firsts = []
    for r in (3, 5):
        for it in range(3):
            b = rng.standard_normal((N, r))
            bt = T.tensor(b, device="cuda")
            x = H.solve(bt)
            xt = H.solve(bt, "T")
            y = H.apply(bt)
            yt = H.apply(bt, "T")
            assert rel(x, np.linalg.solve(A, b)) < 1e-10
            assert rel(xt, np.linalg.solve(A.T, b)) < 1e-10
            assert rel(y, A @ b) < 1e-12
            assert rel(yt, A.T @ b) < 1e-12
            if it == 0:
                firsts.append((x, x.clone()))
    for x, keep in firsts:  # later replays did not write into earlier results
        assert T.equal(x, keep)
    assert len(H._plans) == 8
    monkeypatch.setattr(hbs, "_PLANS", False)
    b = T.tensor(rng.standard_normal((N, 3)), device="cuda")
    x0 = H.solve(b)
    monkeypatch.setattr(hbs, "_PLANS", True)
    assert rel(H.solve(b), x0.cpu().numpy()) < 1e-13


@pytest.mark.parametrize("M,N,K,ta,tb", [(300, 130, 500, 0, 0), (1024, 90, 1024, 1, 0),
                                         (77, 64, 33, 0, 1), (1, 1, 1, 0, 0)])
def test_sgemm(T, M, N, K, ta, tb):
    """FP32 GEMM (the sketch products of the FP32-sketch compression): FP32 accumulation,
    so relative error about sqrt(K) * 2^-24 against the FP64 product."""
    from sslablu_amd.ops import sgemm
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    At = T.tensor(A, device="cuda")
    Bt = T.tensor(B, device="cuda")
    C = sgemm(At.t() if ta else At, Bt.t() if tb else Bt)
    ref = (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64)
    assert C.dtype == T.float32
    assert rel(C.double(), ref) < 1e-5


def test_fp32_sketch_compression(T):
    """Sketches formed in FP32 (HBSMAT(sketch_fp32=True)), compression in FP64: on an
    exactly-HBS operator the compressed operator is as accurate as FP32 sketches allow
    (relative error near 1e-6, where the FP64 sketches give 1e-9) -- the rounding of the
    sketches, not the FP64 pivoting, bounds it."""
    from sslablu_amd.matAssembly.HBS import HBSnew

    class _Tree:  # the oracle generator's identity-ordered binary tree
        def __init__(self, N, nl, L):
            self.perm_leaf = np.arange(N)
            self.nleaves = N // nl
            self.nlevels = L

    rng = np.random.default_rng(3)
    k, nl, L, fac = 6, 16, 4, 2
    U, D, V, Nbvec = oh.gen_random_hbs(k, nl, L, fac, rng)
    A = oh.hbs_dense(U, D, V)
    At = T.tensor(A, device="cuda")
    errs = {}
    for f32 in (False, True):
        h = HBSnew.HBSMAT(At, _Tree(A.shape[0], nl, L), quad=False, seed=4, sketch_fp32=f32)
        h.construct(k)
        x = rng.standard_normal((A.shape[0], 5))
        errs[f32] = rel(h.matmat(T.tensor(x, device="cuda")), A @ x)
        # stage times come from HIP events, resolved when read; compute_ULV adds to tCompress
        ts, tc = h.tSample, h.tCompress
        assert ts > 0 and tc > 0
        h.compute_ULV()
        assert h.tCompress > tc and h.tSample == ts
        # the seeded device generator repeats its draws (split-K atomics may reorder sums)
        h2 = HBSnew.HBSMAT(At, _Tree(A.shape[0], nl, L), quad=False, seed=4, sketch_fp32=f32)
        h2.construct(k)
        assert rel(h2._h.Dmats[-1], h._h.Dmats[-1].cpu().numpy()) < 1e-12
    assert errs[False] < 1e-9
    assert errs[True] < 1e-5


def test_grouped_gemm(T):
    """sslab_dgemm_grouped: independent descriptors with summed terms (transposed operands,
    a broadcast batch-1 operand, an input matrix with beta), more than 12 descriptors in
    one call (split over launches), against numpy."""
    from sslablu_amd.hbs import gdesc, ggemm
    rng = np.random.default_rng(9)
    dev = "cuda"
    specs = []
    for i in range(14):
        nb, M, N = 3, 20 + i, 5 + (i % 4)
        K1, K2 = 7 + i, 33
        A1 = rng.standard_normal((nb, K1, M))          # used transposed
        B1 = rng.standard_normal((nb, K1, N))
        A2 = rng.standard_normal((1, M, K2))            # broadcast over the batch
        B2 = rng.standard_normal((nb, N, K2))           # used transposed
        C0 = rng.standard_normal((nb, M, N))
        ref = 0.5 * C0 + 2.0 * np.swapaxes(A1, 1, 2) @ B1 - A2 @ np.swapaxes(B2, 1, 2)
        t = [T.tensor(x, device=dev) for x in (A1, B1, A2, B2, C0)]
        out = T.empty(nb, M, N, dtype=T.float64, device=dev)
        d = gdesc(out, [(t[0], t[1], 2.0, True, False), (t[2], t[3], -1.0, False, True)],
                  cin=t[4], beta=0.5)
        specs.append((d, out, ref, t))
    ggemm(*[d for d, _, _, _ in specs])
    for _, out, ref, _ in specs:
        assert rel(out, ref) < 1e-13
