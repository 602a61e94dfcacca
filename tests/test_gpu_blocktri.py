"""GPU parity: dense primitives and block-tridiagonal solvers vs the oracle.

Tolerance (plain FP64 GPU path): relative error <= 1e-10 against the FP64 oracle.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import blocktri as ob

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.fixture(scope="module")
def torch_dev(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch, gpu


@pytest.mark.parametrize("M,N,K,ta,tb", [
    (64, 64, 64, 0, 0), (100, 37, 53, 0, 0), (512, 64, 512, 0, 0), (513, 129, 300, 0, 0),
    (77, 91, 45, 1, 0), (77, 91, 45, 0, 1), (130, 33, 17, 1, 1), (1, 1, 1, 0, 0),
    (256, 256, 256, 0, 0), (1024, 1024, 64, 0, 0)])
def test_dgemm(torch_dev, M, N, K, ta, tb):
    torch, dev = torch_dev
    from sslablu_amd.ops import dgemm
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C0 = rng.standard_normal((M, N))
    At = torch.tensor(A, device=dev)
    Bt = torch.tensor(B, device=dev)
    Ct = torch.tensor(C0, device=dev)
    dgemm(At.t() if ta else At, Bt.t() if tb else Bt, Ct, alpha=-0.5, beta=2.0)
    ref = -0.5 * ((A.T if ta else A) @ (B.T if tb else B)) + 2.0 * C0
    assert rel(Ct.cpu().numpy(), ref) < 1e-13


@pytest.mark.parametrize("M,N,K,ldb,batch", [
    (130, 66, 34, 66, 1), (200, 3, 100, 4, 2), (64, 64, 4096, 64, 1), (96, 6, 1538, 8, 3),
    (513, 130, 770, 130, 2), (17, 1, 6, 2, 5)])
def test_dgemm_vector_path_edges(torch_dev, M, N, K, ldb, batch):
    """Even K / leading dimensions and 16-byte aligned operands take the vector-load
    kernel: ragged M, N and K tails, N below the vector width (ldb > N), and long K
    (split over several workgroups)."""
    torch, dev = torch_dev
    from sslablu_amd._lib import call, ptr, stream_handle
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((batch, M, K))
    Bfull = rng.standard_normal((batch, K, ldb))
    C0 = rng.standard_normal((batch, M, N))
    At, Bt, Ct = (torch.tensor(x, device=dev) for x in (A, Bfull, C0))
    call("sslab_dgemm_strided_batched", 0, 0, M, N, K, 0.75, ptr(At), K, M * K, ptr(Bt), ldb,
         K * ldb, -1.5, ptr(Ct), N, M * N, batch, stream_handle())
    ref = 0.75 * (A @ Bfull[:, :, :N]) - 1.5 * C0
    assert rel(Ct.cpu().numpy(), ref) < 1e-13


def test_dgemm_strided_batched(torch_dev):
    torch, dev = torch_dev
    from sslablu_amd._lib import call, ptr, stream_handle
    rng = np.random.default_rng(5)
    b, M, N, K = 7, 96, 48, 80
    A = rng.standard_normal((b, M, K))
    B = rng.standard_normal((b, K, N))
    At, Bt = torch.tensor(A, device=dev), torch.tensor(B, device=dev)
    Ct = torch.zeros(b, M, N, dtype=torch.float64, device=dev)
    call("sslab_dgemm_strided_batched", 0, 0, M, N, K, 1.0, ptr(At), K, M * K, ptr(Bt), N, K * N,
         0.0, ptr(Ct), N, M * N, b, stream_handle())
    assert rel(Ct.cpu().numpy(), A @ B) < 1e-13


@pytest.mark.parametrize("n,batch", [(1, 1), (5, 3), (33, 2), (64, 4), (200, 3), (512, 2),
                                     (1000, 1), (1100, 2), (2600, 1)])
def test_getrf_getrs(torch_dev, n, batch):
    torch, dev = torch_dev
    from sslablu_amd._lib import call, ptr, stream_handle
    rng = np.random.default_rng(n + batch)
    A = rng.standard_normal((batch, n, n))
    B = rng.standard_normal((batch, n, 7))
    At = torch.tensor(A, device=dev)
    piv = torch.empty(batch, n, dtype=torch.int32, device=dev)
    info = torch.zeros(batch, dtype=torch.int32, device=dev)
    s = stream_handle()
    call("sslab_dgetrf_batched", n, ptr(At), n, n * n, ptr(piv), n, ptr(info), batch, s)
    Bt = torch.tensor(B, device=dev)
    Xt = torch.empty_like(Bt)
    perm = torch.empty(batch, n, dtype=torch.int32, device=dev)
    call("sslab_dgetrs_batched", n, 7, ptr(At), n, n * n, ptr(piv), n, ptr(Bt), 7, n * 7,
         ptr(Xt), 7, n * 7, ptr(perm), batch, s)
    X = Xt.cpu().numpy()
    assert int(info.abs().sum()) == 0
    for k in range(batch):
        lu, p = sla.lu_factor(A[k])
        x_ref = np.linalg.solve(A[k], B[k])
        assert rel(X[k], x_ref) < 1e-10
        # same packed factor as LAPACK (pivot choices are well separated here)
        assert np.array_equal(piv[k].cpu().numpy(), p)
        assert rel(At[k].cpu().numpy(), lu) < 1e-12


@pytest.mark.parametrize("n,tie_rows", [
    (48, (5, 20, 40)),           # one wave, 64-column register panel
    (200, (7, 30, 150, 199)),    # ties inside a wave and across waves
    (450, (70, 71, 300, 449)),   # 512-thread panel
    (900, (100, 640, 899)),      # 32-column register panel, 1024 threads
    (1500, (6, 1030, 1499)),     # two rows per lane (rows 6 and 1030 share lane 6)
])
def test_getrf_pivot_ties_first_index(torch_dev, n, tie_rows):
    """Exact ties for the largest |a| in the first column go to the first row (LAPACK's
    idamax); random data elsewhere keeps the later pivots well separated."""
    torch, dev = torch_dev
    from sslablu_amd._lib import call, ptr, stream_handle
    rng = np.random.default_rng(n)
    A = rng.standard_normal((2, n, n))
    A[:, :, 0] = np.clip(A[:, :, 0], -4.0, 4.0)
    for k, r in enumerate(tie_rows):
        A[:, r, 0] = 9.0 if k % 2 == 0 else -9.0
    At = torch.tensor(A, device=dev)
    piv = torch.empty(2, n, dtype=torch.int32, device=dev)
    info = torch.zeros(2, dtype=torch.int32, device=dev)
    call("sslab_dgetrf_batched", n, ptr(At), n, n * n, ptr(piv), n, ptr(info), 2, stream_handle())
    assert int(info.abs().sum()) == 0
    for k in range(2):
        lu, p = sla.lu_factor(A[k])
        got = piv[k].cpu().numpy()
        assert got[0] == min(tie_rows)
        assert np.array_equal(got, p)
        assert rel(At[k].cpu().numpy(), lu) < 1e-12


def test_getrf_reports_singular(torch_dev):
    torch, dev = torch_dev
    from sslablu_amd._lib import call, ptr, stream_handle
    A = np.eye(6)
    A[3, 3] = 0.0
    At = torch.tensor(A, device=dev)
    piv = torch.empty(6, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    call("sslab_dgetrf_batched", 6, ptr(At), 6, 0, ptr(piv), 0, ptr(info), 1, stream_handle())
    assert int(info[0]) == 4  # LAPACK convention: 1-based index of the zero pivot


@pytest.mark.parametrize("nSlabs,m", [(8, 3), (2, 1), (2, 5), (4, 3), (8, 1), (8, 5), (8, 10),
                                      (16, 3), (32, 3)])
@pytest.mark.parametrize("cls", ["RedBlackSolver", "BlockTridiagonalSolver"])
def test_solvers_kat(torch_dev, nSlabs, m, cls):
    """The reference's test_RB known-answer cases (direct_solve/test_RB.py:70-81)."""
    from sslablu_amd.direct_solve import omsdirectsolve as ods
    S, T, dense, rhs = ob.make_block_tridiagonal(nSlabs, m, seed=42)
    solver = getattr(ods, cls)(m, cyclic=False)
    solver.factorize(S, T)
    x = solver.solve(rhs)
    assert isinstance(x, np.ndarray) and x.shape == rhs.shape
    x_ref = np.linalg.solve(dense, rhs)
    assert rel(x, x_ref) < 1e-8
    assert rel(x, x_ref) < TOL


@pytest.mark.parametrize("nb,m,r", [(3, 7, 2), (5, 16, 3), (7, 160, 1), (12, 33, 5),
                                    (31, 128, 8), (31, 512, 64)])
@pytest.mark.parametrize("method", ["rb", "thomas"])
def test_blocktri_vs_oracle(torch_dev, nb, m, r, method):
    torch, dev = torch_dev
    from sslablu_amd.blocktri import BlockTriFactor
    rng = np.random.default_rng(nb * 1000 + m)
    T = [np.eye(m) + 0.3 * rng.standard_normal((m, m)) / np.sqrt(m) for _ in range(nb)]
    S = [(0.5 * rng.standard_normal((m, m)) / np.sqrt(m) if i > 0 else np.zeros((m, m)),
          0.5 * rng.standard_normal((m, m)) / np.sqrt(m) if i < nb - 1 else np.zeros((m, m)))
         for i in range(nb)]
    rhs = rng.standard_normal((nb * m, r))
    fac = BlockTriFactor(method).factorize(np.stack(T), np.stack([s[0] for s in S]),
                                           np.stack([s[1] for s in S]))
    x = fac.solve(torch.tensor(rhs, device=dev)).cpu().numpy()
    x_ref = ob.thomas_direct(S, T, rhs)
    assert rel(x, x_ref) < TOL
    # residual against the assembled operator
    if nb * m <= 4096:
        A = ob.assemble_dense(S, T)
        assert np.linalg.norm(A @ x - rhs) / np.linalg.norm(rhs) < 1e-12


@pytest.mark.parametrize("nb", [2, 4, 5, 8])
@pytest.mark.parametrize("cls", ["RedBlackSolver", "BlockTridiagonalSolver"])
def test_cyclic(torch_dev, nb, cls):
    from sslablu_amd.direct_solve import omsdirectsolve as ods
    m = 6
    rng = np.random.default_rng(nb)
    T = [np.eye(m) * 3 + 0.2 * rng.standard_normal((m, m)) for _ in range(nb)]
    S = [(0.3 * rng.standard_normal((m, m)), 0.3 * rng.standard_normal((m, m))) for _ in range(nb)]
    rhs = rng.standard_normal((nb * m, 2))
    x_ref = np.linalg.solve(ob.assemble_dense(S, T, cyclic=True), rhs)
    solver = getattr(ods, cls)(m, cyclic=True)
    solver.factorize(S, T)
    assert rel(solver.solve(rhs), x_ref) < TOL


def test_thomas_attributes_and_glob_dofs(torch_dev):
    torch, dev = torch_dev
    from sslablu_amd.direct_solve import omsdirectsolve as ods
    S, T, dense, rhs = ob.make_block_tridiagonal(4, 5, seed=42)
    solver = ods.BlockTridiagonalSolver(5)
    solver.factorize(S, T)
    A, B, C = ob.thomas_factor(S, T)
    assert len(solver.A) == 3 and len(solver.B) == 4 and len(solver.C) == 3
    for k in range(3):
        assert rel(solver.A[k].cpu().numpy(), A[k]) < 1e-12
        assert rel(solver.C[k].cpu().numpy(), C[k]) < 1e-15
    for k in range(4):
        Bk = ods.reconstruct_from_lu_factor((solver.B[k][0].cpu().numpy(),
                                             solver.B[k][1].cpu().numpy()))
        assert rel(Bk, ods.reconstruct_from_lu_factor(B[k])) < 1e-12
    dofs = [range(k * 5, (k + 1) * 5) for k in range(4)]
    assert rel(solver.solve(rhs, glob_target_dofs=dofs), np.linalg.solve(dense, rhs)) < TOL


def test_rb_level_api(torch_dev):
    """build_block_RB_solver_level on explicit lists matches the oracle level."""
    torch, dev = torch_dev
    from sslablu_amd.direct_solve import omsdirectsolve as ods
    S, T, dense, rhs = ob.make_block_tridiagonal(4, 3, seed=42)
    RB0 = ob.redblack_factor(S, T)[0]
    ref = ob.redblack_level(3, 4, RB0)
    got = ods.RedBlackSolver(3).build_block_RB_solver_level(3, 4, RB0)
    for k in range(2):
        assert rel(got[1][k].cpu().numpy(), ref[1][k]) < 1e-12
        for j in (0, 3):  # zero boundary couplings compare exactly
            if np.linalg.norm(ref[j][k]) == 0:
                assert np.linalg.norm(got[j][k].cpu().numpy()) <= 1e-12
            else:
                assert rel(got[j][k].cpu().numpy(), ref[j][k]) < 1e-12


@pytest.mark.parametrize("nb", [1, 2, 3, 7, 16])
@pytest.mark.parametrize("method", ["rb", "thomas"])
def test_identity_diagonal_shortcut(torch_dev, nb, method):
    """T = None (identity diagonal blocks, the oms S-system) equals passing T = I."""
    torch, dev = torch_dev
    from sslablu_amd.blocktri import BlockTriFactor
    m, r = 40, 3
    rng = np.random.default_rng(nb)
    SiM = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiP = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiM[0] = 0
    SiP[-1] = 0
    B = torch.tensor(rng.standard_normal((nb * m, r)), device=dev)
    f0 = BlockTriFactor(method).factorize(None, SiM, SiP)
    f1 = BlockTriFactor(method).factorize(np.stack([np.eye(m)] * nb), SiM, SiP)
    x0, x1 = f0.solve(B), f1.solve(B)
    assert int((f0.info != 0).sum()) == 0
    assert (torch.linalg.norm(x0 - x1) / torch.linalg.norm(x1)).item() < 1e-12
    S = [(SiM[i], SiP[i]) for i in range(nb)]
    ref = ob.thomas_direct(S, [np.eye(m)] * nb, B.cpu().numpy())
    assert rel(x0.cpu().numpy(), ref) < TOL


def _rb_plan_shape(nb, m):
    """Levels and coarsest-level size of the red-black factor: at least one level (nb >= 2),
    then halving while more than min(4, 2048 // m) nodes remain."""
    top_max = max(1, min(4, 2048 // m))
    n, levels = nb, []
    while n > 1 and (not levels or n > top_max):
        levels.append((n, (n + 1) // 2, n // 2))
        n = (n + 1) // 2
    return levels, n


@pytest.mark.parametrize("nb,m,ntop", [(2, 16, 1), (3, 16, 2), (5, 16, 3), (7, 16, 4),
                                       (9, 16, 3), (15, 16, 4), (5, 700, 2), (7, 1100, 1),
                                       (13, 512, 4)])
def test_rb_dense_coarsest_level(torch_dev, nb, m, ntop):
    """The coarsest red-black level (up to four nodes) is solved through the explicit
    inverse of its dense system; with four nodes two descriptors add into every block
    atomically.  Checked against the Thomas oracle for general diagonal blocks, for
    T = None and through the FP32 factor with FP64 refinement; the factor size pins the
    plan (levels of [Tinv | Tinv SiM | Tinv SiP] and [W^L | W^R] blocks, then the
    (ntop m)^2 inverse)."""
    torch, dev = torch_dev
    from sslablu_amd import _lib
    from sslablu_amd.blocktri import BlockTriFactor
    levels, n_top = _rb_plan_shape(nb, m)
    assert n_top == ntop
    want = sum(nO * 3 * m * m + nE * 2 * m * m for _, nE, nO in levels) + (n_top * m) ** 2
    assert _lib.sizes("rb", nb, m)[0] == want
    r = 5
    rng = np.random.default_rng(nb * 7 + m)
    T = np.stack([np.eye(m) + 0.3 * rng.standard_normal((m, m)) / np.sqrt(m)
                  for _ in range(nb)])
    SiM = 0.4 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiP = 0.4 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiM[0] = 0
    SiP[-1] = 0
    B = rng.standard_normal((nb * m, r))
    S = [(SiM[i], SiP[i]) for i in range(nb)]
    for Tk, Tl in ((T, list(T)), (None, [np.eye(m)] * nb)):
        fac = BlockTriFactor("rb").factorize(Tk, SiM, SiP)
        assert int((fac.info != 0).sum()) == 0
        x = fac.solve(torch.tensor(B, device=dev)).cpu().numpy()
        ref = ob.thomas_direct(S, Tl, B)
        assert rel(x, ref) < TOL
    fm = BlockTriFactor("rb").factorize_mixed(T, SiM, SiP)
    Bt = torch.tensor(B, device=dev)
    X = torch.empty_like(Bt)
    fm.solve_mixed_(Bt.clone(), X, iters=3)
    assert rel(X.cpu().numpy(), ob.thomas_direct(S, list(T), B)) < 1e-10


@pytest.mark.parametrize("nb,m", [(1, 8), (2, 16), (7, 16), (12, 40), (31, 64)])
@pytest.mark.parametrize("ident", [False, True])
def test_rb_split_phases(torch_dev, nb, m, ident):
    """Forward-only then backward-only equals the whole red-black solve (general and
    identity diagonal blocks); the backward phase reads no V."""
    torch, dev = torch_dev
    from sslablu_amd.blocktri import BlockTriFactor
    rng = np.random.default_rng(nb * 3 + m)
    T = None if ident else np.stack([np.eye(m) + 0.3 * rng.standard_normal((m, m)) / np.sqrt(m)
                                     for _ in range(nb)])
    SiM = 0.4 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiP = 0.4 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiM[0] = 0
    SiP[-1] = 0
    B = torch.tensor(rng.standard_normal((nb * m, 6)), device=dev)
    fac = BlockTriFactor("rb").factorize(T, SiM, SiP)
    X0 = fac.solve(B)
    X1 = torch.empty_like(B)
    fac.solve_(B.clone(), X1, phase="forward")
    fac.solve_(None, X1, phase="backward")
    assert rel(X1.cpu().numpy(), X0.cpu().numpy()) < 1e-13
    with pytest.raises(ValueError):
        BlockTriFactor("thomas").factorize(T, SiM, SiP).solve_(B.clone(), X1, phase="forward")


@pytest.mark.parametrize("bad_level", [0, 1])
def test_rb_factor_recovers_from_a_failed_inverse(torch_dev, bad_level):
    """The red-black factor defers every inverse's backward-error check to one read at the
    end; a diagonal block whose leading 64 x 64 block is zero breaks the block elimination,
    so the whole factorization must be redone with pivoted-LU inverses and still solve the
    system (against the oracle's Thomas solve and the assembled operator).  bad_level 0
    plants the block in an odd node (inverted at level 0); 1 plants it in node 2, which
    reaches level 1's inverses through its Schur complement (node 2's couplings vanish in
    those 64 rows, so the block stays exactly zero there)."""
    torch, dev = torch_dev
    from sslablu_amd.blocktri import BlockTriFactor
    rng = np.random.default_rng(11 + bad_level)
    nb, m, r, h = 7, 128, 3, 64
    T = [np.eye(m) + 0.3 * rng.standard_normal((m, m)) / np.sqrt(m) for _ in range(nb)]
    B = np.eye(h) + 0.3 * rng.standard_normal((h, h)) / np.sqrt(h)
    C = np.eye(h) + 0.3 * rng.standard_normal((h, h)) / np.sqrt(h)
    T[1 if bad_level == 0 else 2] = np.block([[np.zeros((h, h)), B], [C, np.zeros((h, h))]])
    S = [(0.02 * rng.standard_normal((m, m)) / np.sqrt(m) if i > 0 else np.zeros((m, m)),
          0.02 * rng.standard_normal((m, m)) / np.sqrt(m) if i < nb - 1 else np.zeros((m, m)))
         for i in range(nb)]
    if bad_level == 1:
        S[2][0][:h] = 0.0
        S[2][1][:h] = 0.0
    rhs = rng.standard_normal((nb * m, r))
    fac = BlockTriFactor("rb").factorize(np.stack(T), np.stack([s[0] for s in S]),
                                         np.stack([s[1] for s in S]))
    x = fac.solve(torch.tensor(rhs, device=dev)).cpu().numpy()
    x_ref = ob.thomas_direct(S, T, rhs)
    assert rel(x, x_ref) < 1e-11
    A = ob.assemble_dense(S, T)
    assert np.linalg.norm(A @ x - rhs) / np.linalg.norm(rhs) < 1e-12


@pytest.mark.parametrize("nb,m,ncols", [(2, 24, 90), (5, 24, 200), (8, 40, 300), (13, 16, 257)])
def test_rb_forward_windowed_matches_forward_only(torch_dev, nb, m, ncols):
    """sslab_rb_forward_windowed on a V whose block rows vanish outside random column
    intervals gives the entry state of the dense forward-only sweep (identity diagonal
    blocks) on the same V."""
    import ctypes
    torch, dev = torch_dev
    from sslablu_amd.blocktri import BlockTriFactor
    from sslablu_amd._lib import call, ptr, stream_handle
    rng = np.random.default_rng(nb * 7 + m)
    S = [(0.5 * rng.standard_normal((m, m)) / np.sqrt(m), 0.5 * rng.standard_normal((m, m)) / np.sqrt(m))
         for _ in range(nb)]
    fac = BlockTriFactor("rb").factorize(None, np.stack([s[0] for s in S]),
                                         np.stack([s[1] for s in S]))
    lo = rng.integers(0, ncols - 1, nb)
    hi = np.minimum(ncols, lo + 1 + rng.integers(0, ncols // 3, nb))
    V = np.zeros((nb * m, ncols))
    for i in range(nb):
        V[i * m:(i + 1) * m, lo[i]:hi[i]] = rng.standard_normal((m, hi[i] - lo[i]))
    V0 = torch.tensor(V, device=dev)
    X0 = torch.empty_like(V0)
    call("sslab_rb_solve_ex", nb, m, ncols, ptr(fac.factor), ptr(V0), ncols, ptr(X0), ncols,
         1 | 4, stream_handle())
    V1 = torch.tensor(V, device=dev)
    X1 = torch.full_like(V1, float("nan"))
    lo_a = (ctypes.c_int * nb)(*lo.tolist())
    hi_a = (ctypes.c_int * nb)(*hi.tolist())
    call("sslab_rb_forward_windowed", nb, m, ncols, ptr(fac.factor), ptr(V1), ncols, ptr(X1),
         ncols, ctypes.addressof(lo_a), ctypes.addressof(hi_a), stream_handle())
    torch.cuda.synchronize()
    assert torch.isfinite(X1).all()
    assert (torch.linalg.norm(X1 - X0) / torch.linalg.norm(X0)).item() < 1e-13
