"""Device GMRES (sslablu_amd.krylov): the column-batched vector kernels against plain
FP64 torch arithmetic, GMRES on block-tridiagonal systems without and with an FP32 direct
preconditioner against the FP64 direct solve, and GMRES on the exact 3-D S-system with a
low-rank HBS direct solver as preconditioner (the reference's closing step,
cubeTest.py:163-186) against the dense FP64 solve."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("n,r,k", [(1, 1, 1), (300, 3, 4), (1000, 130, 2), (257, 64, 7)])
def test_col_kernels(T, n, r, k):
    from sslablu_amd import krylov
    from sslablu_amd._lib import stream_handle
    g = T.Generator(device="cuda").manual_seed(n + r + k)
    V = T.randn(k, n, r, dtype=T.float64, device="cuda", generator=g)
    W = T.randn(n, r, dtype=T.float64, device="cuda", generator=g)
    s = stream_handle()
    H = krylov._dots(V, k, W, s)
    assert rel(H.cpu(), (V * W[None]).sum(1).cpu()) < 1e-13
    W2 = W.clone()
    krylov._axpy(W2, -0.5, V, k, H, s)
    assert rel(W2.cpu(), (W - 0.5 * (V * H[:, None, :]).sum(0)).cpu()) < 1e-13
    sc = T.randn(r, dtype=T.float64, device="cuda", generator=g)
    out = T.empty_like(W)
    krylov._scale(W, sc, out, s)
    assert rel(out.cpu(), (W * sc[None]).cpu()) < 1e-15
    assert rel(krylov._norms(W, s), W.norm(dim=0).cpu().numpy()) < 1e-13


def test_gmres_blocktri(T):
    """Well-conditioned Gaussian-kernel block-tridiagonal system (T != I): plain GMRES
    and GMRES with the FP32 red-black factor as preconditioner both reach the FP64
    direct solution; the preconditioned run needs only a few steps."""
    from sslablu_amd.blocktri import BlockTriFactor
    from sslablu_amd.krylov import gmres, tridiag_apply
    from sslablu_amd.synthetic import gaussian_blocktri
    nb, n, r = 6, 64, 5
    Tb, SiM, SiP = gaussian_blocktri(nb, n, seed=2)
    B = T.randn(nb * n, r, dtype=T.float64, device="cuda",
                generator=T.Generator(device="cuda").manual_seed(7))
    x_ref = BlockTriFactor("rb").factorize(Tb, SiM, SiP).solve(B)
    A = tridiag_apply(SiM, SiP, nb, n, Tb)
    assert rel(A(x_ref).cpu(), B.cpu()) < 1e-13
    X, info = gmres(A, B, tol=1e-12, restart=40, maxiter=200)
    assert info.converged and rel(X.cpu(), x_ref.cpu()) < 1e-10
    f32 = BlockTriFactor("rb").factorize_mixed(Tb, SiM, SiP)
    M = lambda v: f32.solve_mixed_(v.contiguous(), T.empty_like(v), iters=0)  # noqa: E731
    X2, info2 = gmres(A, B, apply_M=M, tol=1e-12, restart=10, maxiter=20)
    assert info2.converged and info2.iterations <= 6, info2
    assert rel(X2.cpu(), x_ref.cpu()) < 1e-11


def test_gmres_hbs_preconditioned_3d(T):
    """3-D spectral S-system: a rank-8 HBS Thomas solver alone is far from the dense
    solution; as GMRES preconditioner on the exact operator it reaches 1e-10."""
    from sslablu_amd.blocktri import BlockTriFactor
    from sslablu_amd.direct_solve.omsdirectsolveHBS import ThomasSolverHBS
    from sslablu_amd.krylov import gmres, tridiag_apply
    from sslablu_amd.matAssembly.HBS import HBSnew
    from sslablu_amd.spectral3d import FaceTree, SpectralSlabs3D
    s = SpectralSlabs3D(8, 9, 34, 34).build()
    tree = FaceTree(32, 32, 64)
    rk = 8
    H = {}
    for side in ("left", "right"):
        h = HBSnew.HBSMAT(s.S[side], tree, quad=False, seed=1,
                          samples=max(2 * rk, 64) + rk + 10)
        h.construct(rk)
        H[side] = h
    nb = s.nb
    S_rk = [[H["right"]]] + [[H["left"], H["right"]] for _ in range(nb - 2)] + [[H["left"]]]
    solver = ThomasSolverHBS(s.m, rk)
    solver.factorize(S_rk)
    rng = np.random.default_rng(0)
    B = T.tensor(rng.standard_normal((nb * s.m, 3)), device="cuda")
    SiM, SiP = s.blocks()
    x_dense = BlockTriFactor("rb").factorize(None, SiM, SiP).solve(B)
    x_hbs = solver.solve(B)
    e_hbs = rel(T.as_tensor(x_hbs).cpu(), x_dense.cpu())
    assert e_hbs > 1e-6  # the compressed solver alone is visibly inexact here
    A = tridiag_apply(SiM, SiP, nb, s.m)
    M = lambda v: T.as_tensor(solver.solve(v), device="cuda").contiguous()  # noqa: E731
    X, info = gmres(A, B, apply_M=M, tol=1e-11, restart=20, maxiter=60)
    assert info.converged
    assert rel(X.cpu(), x_dense.cpu()) < 1e-9, (e_hbs, info.history)
