"""GPU parity of the 3-D HPS slab assembly (sslablu_amd/hps3d.py) against the 3-D oracle
(oracle/hps3d.py): leaf DtN maps on harmonic polynomials, S blocks and right-hand sides
(1e-10), the S-system solution (exact on a harmonic polynomial, and equal to the
oracle's global solve), transposed applies through the transposed LU solve, and the
batched transposed solve itself against numpy.

Tolerances: 1e-10 relative for S blocks, rhs and solutions (FP64 throughout); 1e-12
for the LU solves."""
import numpy as np
import pytest

from oracle import hps3d as o3

pytestmark = pytest.mark.gpu


class PDO3:
    def __init__(self, c=None, c1=None):
        one = lambda P: np.ones(P.shape[0])  # noqa: E731
        self.c11 = self.c22 = self.c33 = one
        self.c1, self.c2, self.c3, self.c = c1, None, None, c


def poly(P):
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    return x ** 2 + y ** 2 - 2 * z ** 2 + x * y * z + 3 * x - y


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


@pytest.mark.parametrize("n,nrhs", [(37, 5), (300, 17), (600, 3)])
def test_getrs_trans_batched(T, n, nrhs):
    from sslablu_amd._lib import call, ptr
    rng = np.random.default_rng(n)
    A = rng.standard_normal((3, n, n)) + 0.1 * n ** 0.5 * np.eye(n)
    B = rng.standard_normal((3, n, nrhs))
    dA = T.tensor(A, device="cuda")
    piv = T.empty(3, n, dtype=T.int32, device="cuda")
    call("sslab_dgetrf_batched", n, ptr(dA), n, n * n, ptr(piv), n, None, 3, None)
    dB = T.tensor(B, device="cuda")
    X = T.empty_like(dB)
    W = T.empty_like(dB)
    perm = T.empty(3, n, dtype=T.int32, device="cuda")
    call("sslab_dgetrs_trans_batched", n, nrhs, ptr(dA), n, n * n, ptr(piv), n, ptr(dB), nrhs,
         n * nrhs, ptr(X), nrhs, n * nrhs, ptr(W), ptr(perm), 3, None)
    # Forward error bound of a backward-stable solve, cond(A) eps with a margin: the
    # trailing updates add split-K partial sums atomically, so the rounding (not the
    # accuracy) varies from run to run (random A of size 600 here: cond ~ 1e3).
    for b in range(3):
        tol = max(1e-12, 50 * np.linalg.cond(A[b]) * np.finfo(float).eps)
        assert rel(X[b], np.linalg.solve(A[b].T, B[b])) < tol


@pytest.mark.parametrize("q", [5, 8])
def test_leaf_dtn_exact_on_harmonic(T, q):
    """One leaf: T applied to the face values of a harmonic polynomial (degree <= q - 1
    per variable) gives its outward normal derivative."""
    from sslablu_amd import hps3d as h3
    g = h3.SlabGrid3D(2, 1, 1, 1, q, size=(0.6, 0.25, 0.4))
    op = h3.HPSSlabOperator3D(PDO3(), g)
    Tl = op.leaf_dtn(range(0, 1))[0].cpu().numpy()
    h = (g.hx, g.hy, g.hz)
    t = h3.cheb_unit(q)[0]
    K = h3.box_nodes((1, 1, 1), q)
    X = np.stack([t[K[:, d]] * h[d] for d in range(3)], 1)
    u = poly(X)
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    grad = np.stack([2 * x + y * z + 3, 2 * y + x * z - 1, -4 * z + x * y], 1)
    p2 = (q - 2) ** 2
    want = np.concatenate([(-1 if f % 2 == 0 else 1) * grad[f * p2:(f + 1) * p2, f // 2]
                           for f in range(6)])
    assert rel(Tl @ u, want) < 1e-10


@pytest.mark.parametrize("N,q,shape,pdo", [(3, 5, (1, 2, 2), "laplace"),
                                           (4, 6, (1, 2, 1), "helmholtz"),
                                           (3, 5, (2, 2, 2), "convection")])
def test_s_blocks_and_rhs_vs_oracle(T, N, q, shape, pdo):
    from sslablu_amd import hps3d as h3
    if pdo == "laplace":
        P = PDO3()
    elif pdo == "helmholtz":
        P = PDO3(c=lambda X: -25.0 * (1.0 - 0.5 * np.exp(-((X - 0.5) ** 2).sum(1) / 0.05)))
    else:
        P = PDO3(c1=lambda X: 1.0 + X[:, 1])
    nx, ny, nz = shape
    g = h3.SlabGrid3D(N, nx, ny, nz, q)
    op = h3.HPSSlabOperator3D(P, g).build()
    res = o3.oms_stot(P, N, (g.hx, g.hy, g.hz), q, bc=poly)
    SiM, SiP = op.dense_blocks()
    S = res["S_rk_list"]
    nb = g.nb
    for i in range(nb):
        if i > 0:
            assert rel(SiM[i], S[i][0]) < 1e-10
        if i < nb - 1:
            assert rel(SiP[i], S[i][-1]) < 1e-10
    F = T.tensor(poly(g.gb_points()), device="cuda")
    rhs = op.rhs_from_boundary(F)
    assert rel(rhs[:, 0], np.concatenate(res["rhs_list"])) < 1e-10
    # the operator forms: smap / transposes against the dense blocks
    rng = np.random.default_rng(1)
    V = T.tensor(rng.standard_normal((g.m, 3)), device="cuda")
    for i, side, D in ((nb - 1, "l", SiM), (0, "r", SiP)):
        Sm = op.smap(i, side)
        assert rel(Sm.matmat(V), (D[i] @ V).cpu().numpy()) < 1e-12
        assert rel(Sm.rmatmat(V), (D[i].t() @ V).cpu().numpy()) < 1e-12
    Vb = T.tensor(rng.standard_normal((nb, g.m, 2)), device="cuda")
    assert rel(op.smaps_apply("l", Vb, transpose=True),
               T.bmm(SiM.transpose(1, 2).contiguous(), Vb).cpu().numpy()) < 1e-12


def test_solution_exact_and_equal_to_global_solve(T):
    """The S-system of the GPU operator reproduces a harmonic polynomial exactly at the
    interface nodes and equals the oracle's single global HPS solve (exact domain
    decomposition)."""
    from sslablu_amd import hps3d as h3
    from sslablu_amd.blocktri import BlockTriFactor
    N, q = 4, 6
    g = h3.SlabGrid3D(N, 1, 2, 2, q)
    op = h3.HPSSlabOperator3D(PDO3(), g).build()
    SiM, SiP = op.dense_blocks()
    F = T.tensor(poly(g.gb_points()), device="cuda")
    x = BlockTriFactor("rb").factorize(None, SiM, SiP).solve(op.rhs_from_boundary(F))
    pts = g.interface_points()
    assert rel(x[:, 0], poly(pts)) < 1e-10
    e = lambda P: np.exp(np.sqrt(2) * np.pi * P[:, 0]) * np.sin(np.pi * P[:, 1]) * np.sin(np.pi * P[:, 2])  # noqa: E731
    Fe = T.tensor(e(g.gb_points()), device="cuda")
    xe = BlockTriFactor("rb").factorize(None, SiM, SiP).solve(op.rhs_from_boundary(Fe))
    ug, pg = o3.global_solution_on_planes(PDO3(), N, (g.hx, g.hy, g.hz), q, e)
    assert np.allclose(pg, pts, atol=1e-15)
    assert rel(xe[:, 0], ug) < 1e-10
    # the exact operator apply (GMRES's operator) on the solution gives back the rhs
    A = op.tridiag_apply()
    r = op.rhs_from_boundary(Fe)
    assert rel(A(xe), r.cpu().numpy()) < 1e-10


def test_c3_full_size_hps_hbs_gmres(T):
    """The C3 configuration at full size: 16 slabs of 1 x 16 x 16 leaves (q = 8, m = 9216,
    15 interfaces), every S block HBS-compressed at rank 40 (leaf 72, s = 130, FP32
    sketch GEMMs), ULV-factored HBS Thomas elimination as the preconditioner of GMRES on
    the exact S-system, 16 right-hand sides.  Size-independent checks: every column's
    recomputed relative residual ||b - S x|| / ||b|| <= 1e-10, and the column with a
    harmonic polynomial's boundary data reproduces the polynomial at all 138240
    interface nodes (the q = 8 leaves are exact for it)."""
    import torch
    from sslablu_amd import hps3d as h3
    from sslablu_amd.direct_solve.omsdirectsolveHBS import ThomasSolverHBS
    from sslablu_amd.krylov import gmres
    from sslablu_amd.spectral3d import FaceTree
    print("[c3] full size: 16 slabs, q=8, 16x16 leaves per interface", flush=True)
    g = h3.SlabGrid3D(16, 1, 16, 16, 8)
    op = h3.HPSSlabOperator3D(PDO3(), g, slab_chunk=4).build()
    assert op.info_bad == 0 and g.m == 9216 and g.nb == 15
    rng = np.random.default_rng(3)
    modes = rng.integers(1, 4, size=(15, 8, 3))
    amps = rng.standard_normal((15, 8))
    P = g.gb_points()
    F = np.concatenate([poly(P)[:, None]] + [
        (amps[c][None, :] * np.sin(np.pi * P @ modes[c].T)).sum(1)[:, None] for c in range(15)],
        axis=1)
    B = op.rhs_from_boundary(T.tensor(F, device="cuda"))
    H_l, H_r = h3.compress_s_blocks(op, FaceTree(96, 96, 72), 40, 130, seed=0, sketch_fp32=True)
    solver = ThomasSolverHBS(g.m, 40, samples=130)
    solver.factorize(h3.hbs_S_rk_list(H_l, H_r))
    A = op.tridiag_apply()
    X, info = gmres(A, B, apply_M=lambda v: solver.solve(v).contiguous(), tol=1e-10,
                    restart=60, maxiter=120)
    print(f"[c3] gmres iterations {info.iterations}", flush=True)
    res = ((A(X) - B).norm(dim=0) / B.norm(dim=0)).cpu().numpy()
    assert np.all(res <= 1e-10), res
    assert rel(X[:, 0], poly(g.interface_points())) < 1e-9
    del op, H_l, H_r, solver
    torch.cuda.empty_cache()
