"""Reference-API drop-in on the GPU: drivers written against SimonDirckx/SslabLU's
modules (geometry.geom_2D.square, pdo.pdo, solver.solver, multislab.oms,
matAssembly.matAssembler, direct_solve.*, multislab.omsdirectsolve) run unchanged on the
sslablu_amd mirrors.  Checked against the exact harmonic solution of config 1 and
against the dense S-system solve.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def exp_sin(P):
    return np.exp(P[:, 0]) * np.sin(P[:, 1])


@pytest.fixture(scope="module")
def c1(gpu):
    from sslablu_amd.geometry.geom_2D import square
    from sslablu_amd.matAssembly.matAssembler import denseMatAssembler
    from sslablu_amd.multislab.oms import oms
    from sslablu_amd.pdo.pdo import PDO2d, const
    from sslablu_amd.solver.solver import solverOptions
    dSlabs, conn, H = square.dSlabs(8)
    opts = solverOptions("hps", [12, 12], a=[1 / 32, 1 / 32])
    OMS = oms(dSlabs, PDO2d(c11=const(1), c22=const(1)), square.gb_np, opts, conn)
    S_rk_list, rhs_list, Ntot, nc = OMS.construct_Stot_helper(exp_sin, denseMatAssembler())
    return OMS, S_rk_list, rhs_list, Ntot, nc


def test_oms_shapes_and_exact_solution(c1):
    from sslablu_amd.direct_solve.omsdirectsolve import BlockTridiagonalSolver, RedBlackSolver
    OMS, S_rk_list, rhs_list, Ntot, nc = c1
    assert nc == 160 and Ntot == 7 * 160 and len(S_rk_list) == 7
    assert len(S_rk_list[0]) == 1 and len(S_rk_list[3]) == 2 and len(S_rk_list[-1]) == 1
    Stot, rhstot = OMS.construct_Stot_and_rhstot_linearOperator(S_rk_list, rhs_list, Ntot, nc)
    pts = OMS.discretize().g.interface_points()
    T = [np.eye(nc)] * len(S_rk_list)
    for cls in (BlockTridiagonalSolver, RedBlackSolver):
        solver = cls(nc)
        solver.factorize(S_rk_list, T)
        x = solver.solve(rhstot)
        assert np.max(np.abs(x - exp_sin(pts))) < 1e-10
        assert np.linalg.norm(Stot @ x - rhstot) / np.linalg.norm(rhstot) < 1e-12
    # S blocks answer numpy products like the reference drivers use (cubeTest.py:150-155)
    E = np.identity(nc)
    dense = S_rk_list[2][0] @ E
    assert isinstance(dense, np.ndarray) and dense.shape == (nc, nc)


def test_old_identity_diagonal_api(c1):
    from sslablu_amd.multislab import omsdirectsolve as old
    OMS, S_rk_list, rhs_list, Ntot, nc = c1
    _, rhstot = OMS.construct_Stot_and_rhstot_linearOperator(S_rk_list, rhs_list, Ntot, nc)
    fac = old.build_block_tridiagonal_solver(S_rk_list)
    x = old.block_tridiagonal_solve(OMS, fac, rhstot)
    pts = OMS.discretize().g.interface_points()
    assert np.max(np.abs(x - exp_sin(pts))) < 1e-10


def test_hbs_assembler_and_thomas_hbs(gpu):
    """rkHBS assembly (binary slabTree, leaf 40, rank 20) of the config-1 S blocks and
    the HBS Thomas solver: agrees with the dense solution to the compression accuracy."""
    from sslablu_amd.direct_solve.omsdirectsolveHBS import ThomasSolverHBS
    from sslablu_amd.geometry.geom_2D import square
    from sslablu_amd.matAssembly.matAssembler import rkHMatAssembler
    from sslablu_amd.multislab.oms import oms
    from sslablu_amd.pdo.pdo import PDO2d, const
    from sslablu_amd.solver.solver import solverOptions
    dSlabs, conn, H = square.dSlabs(8)
    opts = solverOptions("hps", [12, 12], a=[1 / 32, 1 / 32])
    OMS = oms(dSlabs, PDO2d(c11=const(1), c22=const(1)), square.gb_np, opts, conn)
    S_rk_list, rhs_list, Ntot, nc = OMS.construct_Stot_helper(exp_sin, rkHMatAssembler(40, 20))
    assert OMS.stats.compression < 1.0
    Stot, rhstot = OMS.construct_Stot_and_rhstot_linearOperator(S_rk_list, rhs_list, Ntot, nc)
    # compressed blocks reproduce the dense ones
    S_dense = OMS.discretize().SiP[0].cpu().numpy()
    Hb = S_rk_list[0][0]
    err = np.linalg.norm(np.asarray(Hb @ np.eye(nc)) - S_dense) / np.linalg.norm(S_dense)
    assert err < 1e-6
    solver = ThomasSolverHBS(nc, 20)
    solver.factorize(S_rk_list)
    x = solver.solve(rhstot)
    pts = OMS.discretize().g.interface_points()
    assert np.max(np.abs(x - exp_sin(pts))) < 1e-5
    assert np.linalg.norm(Stot @ x - rhstot) / np.linalg.norm(rhstot) < 1e-5
    # the older identity-diagonal API (multislab/omsdirectsolveHBS.py) on the same blocks
    from sslablu_amd.multislab import omsdirectsolveHBS as old_hbs
    tree = S_rk_list[0][0].tree
    Tf = old_hbs.build_block_tridiagonal_solver(S_rk_list, tree, S_rk_list[0][0].quad, 20)
    x2 = old_hbs.block_tridiagonal_solve(OMS, Tf, rhstot)
    assert np.linalg.norm(x2 - x) / np.linalg.norm(x) < 1e-8
    x2m = old_hbs.block_tridiagonal_solve(OMS, Tf, np.stack([rhstot, 2 * rhstot], 1))
    assert np.linalg.norm(x2m[:, 1] - 2 * x2) / np.linalg.norm(x2) < 1e-10


def test_redblack_hbs(gpu):
    from sslablu_amd.direct_solve.omsdirectsolve import RedBlackSolver
    from sslablu_amd.direct_solve.omsdirectsolveHBS import RedBlackSolverHBS
    from sslablu_amd.geometry.geom_2D import square
    from sslablu_amd.matAssembly.matAssembler import rkHMatAssembler
    from sslablu_amd.multislab.oms import oms
    from sslablu_amd.pdo.pdo import PDO2d, const
    from sslablu_amd.solver.solver import solverOptions
    dSlabs, conn, H = square.dSlabs(9)          # 8 double slabs (a power of two)
    opts = solverOptions("hps", [10, 10], a=[1 / 36, 1 / 32])
    OMS = oms(dSlabs, PDO2d(c11=const(1), c22=const(1)), square.gb_np, opts, conn)
    S_rk_list, rhs_list, Ntot, nc = OMS.construct_Stot_helper(exp_sin, rkHMatAssembler(32, 16))
    _, rhstot = OMS.construct_Stot_and_rhstot_linearOperator(S_rk_list, rhs_list, Ntot, nc)
    rb = RedBlackSolverHBS(nc, 16, S_rk_list[0][0].tree, S_rk_list[0][0].quad)
    rb.factorize(S_rk_list)
    x = rb.solve(rhstot)
    ref = RedBlackSolver(nc)
    ref.factorize(S_rk_list, [np.eye(nc)] * len(S_rk_list))
    x_ref = ref.solve(rhstot)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-5


def test_hbstorch_from_caller_sketches(gpu):
    """HBStorch.HBSMAT.construct(rk, Om, Psi, Y, Z) with sketches in the original
    (unpermuted) order reproduces an exactly-HBS operator behind a leaf permutation,
    and its ULV solve inverts it (HBStorch.py:163-241, 467-493)."""
    from oracle import hbs as oh
    from sslablu_amd.matAssembly.HBS import HBStorch
    rng = np.random.default_rng(21)
    k, nl, L, fac = 5, 12, 4, 2
    U, D, V, Nbvec = oh.gen_random_hbs(k, nl, L, fac, rng)
    D[0] = D[0] + 10 * np.eye(D[0].shape[1])
    H = oh.hbs_dense(U, D, V)
    N = H.shape[0]
    perm = rng.permutation(N)
    A = np.empty_like(H)
    A[np.ix_(perm, perm)] = H  # leaf order = rows perm of the original order

    class _Tree:
        perm_leaf = perm
        nleaves = Nbvec[0]
        nlevels = L

    s = 4 * k + 2 * nl + 5
    Om, Psi = rng.standard_normal((N, s)), rng.standard_normal((N, s))
    hb = HBStorch.HBSMAT(device="cuda", tree=_Tree(), quad=False)
    hb.construct(k, Om, Psi, A @ Om, A.T @ Psi)
    x = rng.standard_normal((N, 3))
    assert np.linalg.norm(hb.matmat(x) - A @ x) / np.linalg.norm(A @ x) < 1e-9
    assert np.linalg.norm(hb.T @ x - A.T @ x) / np.linalg.norm(A.T @ x) < 1e-9
    hb.compute_ULV()
    b = rng.standard_normal(N)
    xs = hb.solve(b, "cuda")
    assert xs.shape == (N,)
    assert np.linalg.norm(A @ xs - b) / np.linalg.norm(b) < 1e-8
    xt = hb.solve(b, "cuda", mode="T")
    assert np.linalg.norm(A.T @ xt - b) / np.linalg.norm(b) < 1e-8


def test_ulvsparse_torch_tensor_form(gpu):
    """ULVsparse_torch.compute_ULV / solve on (Nb, n, k) tensors (ULVsparse_torch.py:231-363)."""
    import torch
    from oracle import hbs as oh
    from sslablu_amd.matAssembly.HBS import ULVsparse_torch
    rng = np.random.default_rng(3)
    U, D, V, Nbvec = oh.gen_random_hbs(4, 8, 3, 2, rng)
    D[0] = D[0] + 10 * np.eye(D[0].shape[1])
    A = oh.hbs_dense(U, D, V)
    t = lambda Xs: [torch.tensor(X, device="cuda") for X in Xs]
    Q, W, Uu, R, NN = ULVsparse_torch.compute_ULV(t(U), t(D), t(V), Nbvec, "cuda")
    b = rng.standard_normal((A.shape[0], 2))
    x = ULVsparse_torch.solve(U, D, Q, W, Uu, R, NN, b, "cuda")
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(b) < 1e-10


def test_ulvsparse_factor_lifetime(gpu):
    """ULVsparse.compute_ULV's device factorization lives exactly as long as the returned
    lists: dropping them releases it, and solve() refuses lists it did not return."""
    import gc
    import weakref

    import pytest
    import torch
    from oracle import hbs as oh
    from sslablu_amd.matAssembly.HBS import ULVsparse
    rng = np.random.default_rng(4)
    U, D, V, Nbvec = oh.gen_random_hbs(4, 8, 3, 2, rng)
    D[0] = D[0] + 10 * np.eye(D[0].shape[1])
    A = oh.hbs_dense(U, D, V)
    Um, Dm, Vm = ([x.reshape(-1, x.shape[-1]) for x in X] for X in (U, D, V))
    Q, W, Uu, R, NN = ULVsparse.compute_ULV(Um, Dm, Vm, Nbvec)
    b = rng.standard_normal(A.shape[0])
    x = ULVsparse.solve(Um, Dm, Q, W, Uu, R, NN, Nbvec, b)
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(b) < 1e-10
    foreign = [q.clone() for q in Q]
    with pytest.raises(ValueError):
        ULVsparse.solve(Um, Dm, foreign, W, Uu, R, NN, Nbvec, b)
    ref = weakref.ref(getattr(Q[0], ULVsparse._ATTR))
    del Q, W, Uu, R, NN, x
    gc.collect()
    assert ref() is None
    torch.cuda.synchronize()
