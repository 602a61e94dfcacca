"""GPU: the HBS block solvers and their operator algebra (direct_solve/omsdirectsolveHBS.py
and multislab/omsdirectsolve.py mirrors) against dense closed forms and the oracle.

* RedBlackSolverHBS at block counts that are not powers of two (and one that is):
  against the oracle's dense FP64 solve of its own oms S-system (the same points) and
  against the exact solution u = exp(x) sin(y) of config 1; the HBS rank is high
  enough that the compression is accurate to ~1e-10, so the tolerance is 1e-8.
* STS_linop, RB_linop, Sprime_Linop, Dprime_Linop (products and transposed products)
  against the dense matrices they stand for (1e-12).
* ThomasSolverHBS.factorize_with_diag (T != I) with one-leaf trees (HBS = dense, exact)
  against the oracle's Thomas solve (1e-10).
* build_block_cyclic_tridiagonal_solver / block_cyclic_tridiagonal_solve against
  numpy.linalg.solve of the assembled periodic matrix (1e-10)."""
import numpy as np
import pytest

from oracle import blocktri as ob
from oracle import hps as oh

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def exp_sin(P):
    return np.exp(P[:, 0]) * np.sin(P[:, 1])


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


@pytest.mark.parametrize("N", [6, 7, 9, 12])
def test_redblack_hbs_any_block_count(T, N):
    from sslablu_amd.direct_solve.omsdirectsolveHBS import RedBlackSolverHBS
    from sslablu_amd.geometry.geom_2D import square
    from sslablu_amd.matAssembly.matAssembler import rkHMatAssembler
    from sslablu_amd.multislab.oms import oms
    from sslablu_amd.pdo.pdo import PDO2d, const
    from sslablu_amd.solver.solver import solverOptions
    q, ny = 12, 16                                  # m = 16 * 10 = 160
    dSlabs, conn, H = square.dSlabs(N)
    opts = solverOptions("hps", [q, q], a=[H / 2, 1 / (2 * ny)])
    pdo = PDO2d(c11=const(1), c22=const(1))
    OMS = oms(dSlabs, pdo, square.gb_np, opts, conn)
    S_rk_list, rhs_list, Ntot, nc = OMS.construct_Stot_helper(exp_sin, rkHMatAssembler(40, 40))
    _, rhstot = OMS.construct_Stot_and_rhstot_linearOperator(S_rk_list, rhs_list, Ntot, nc)
    rb = RedBlackSolverHBS(nc, 40, S_rk_list[0][0].tree, S_rk_list[0][0].quad)
    rb.factorize(S_rk_list)
    assert len(rb.RB[-1][0]) == 1
    x = rb.solve(rhstot)
    ref = oh.oms_stot(pdo, N, (H, 1 / ny), q, bc=exp_sin)
    u = np.linalg.solve(oh.stot_dense(ref["S_rk_list"]), np.concatenate(ref["rhs_list"]))
    assert rel(x, u) < 1e-8
    assert np.max(np.abs(x - exp_sin(np.concatenate(ref["Ic_pts"])))) < 1e-7
    X2 = rb.solve(np.stack([rhstot, 2 * rhstot], 1))      # (N, r) right-hand sides
    assert rel(X2[:, 1], 2 * u) < 1e-8


def _dense_ops(T, m, seed):
    from sslablu_amd.direct_solve import omsdirectsolveHBS as H
    rng = np.random.default_rng(seed)
    mats = [rng.standard_normal((m, m)) / np.sqrt(m) for _ in range(6)]
    mats[1] = mats[1] + 3 * np.eye(m)                    # invertible "T" blocks
    mats[4] = mats[4] + 3 * np.eye(m)
    ops = [H.dense_to_linop(T.tensor(a, device="cuda")) for a in mats]
    return mats, ops


def test_linop_algebra_vs_dense(T):
    from sslablu_amd.direct_solve import omsdirectsolveHBS as H
    m = 48
    (Sl, Tm, Sr, Sp, Tp, D), (oSl, oTm, oSr, oSp, oTp, oD) = _dense_ops(T, m, 4)
    X = np.random.default_rng(5).standard_normal((m, 3))
    inv = np.linalg.inv
    cases = [
        (H.STS_linop(oSl, oTm, oSr), -Sl @ inv(Tm) @ Sr),
        (H.RB_linop(oD, oTm, oTp, oSp, oSl, oSr, oSr), D - Sp @ inv(Tp) @ Sr - Sl @ inv(Tm) @ Sr),
        (H.RB_linop(oD, None, oTp, oSp, oSl, oSr, None), D - Sp @ inv(Tp) @ Sr),
        (H.Sprime_Linop(oSl, oTm, oSr), np.eye(m) - Sl @ inv(Tm) @ Sr),
        (H.Sprime_Linop(oSl, None, oSr, id=True), np.eye(m) - Sl @ Sr),
        (H.Dprime_Linop(oD, oSl, oSr, oTm), D - Sl @ inv(Tm) @ Sr),
    ]
    for op, dense in cases:
        assert rel(op.matmat(X), dense @ X) < 1e-12
        assert rel(op.rmatmat(X), dense.T @ X) < 1e-12
        assert rel(op.T.matmat(X), dense.T @ X) < 1e-12
    idm = H.id_op(m)
    assert rel(idm.matmat(X), X) == 0.0 and rel(idm.solve(X), X) == 0.0
    assert rel(oTm.solve(X), np.linalg.solve(Tm, X)) < 1e-12
    assert rel(oTm.solve(X, "T"), np.linalg.solve(Tm.T, X)) < 1e-12


def test_thomas_hbs_factorize_with_diag(T):
    """T != I through ThomasSolverHBS.factorize_with_diag: every block on a one-leaf tree
    (the HBS form of a block is then the block itself), so the elimination is exact."""
    from sslablu_amd.direct_solve.omsdirectsolveHBS import ThomasSolverHBS
    from sslablu_amd.linop import DeviceMatrix
    from sslablu_amd.matAssembly.HBS.slabTree import slabTree
    nb, m = 5, 40
    S, Tb, dense, rhs = ob.make_block_tridiagonal(nb, m, seed=7)
    tree = slabTree(np.linspace(0, 1, m), min_leaf_size=m)
    assert tree.nleaves == 1

    def dm(a):
        d = DeviceMatrix(T.tensor(a, device="cuda"))
        d.tree, d.quad = tree, False
        return d

    AB = [[dm(S[0][-1])]] + [[dm(S[i][0]), dm(S[i][-1])] for i in range(1, nb - 1)] + \
         [[dm(S[nb - 1][0])]]
    D = [dm(t) for t in Tb]
    solver = ThomasSolverHBS(m, m, samples=m + 10)
    solver.factorize(AB, D)
    assert solver.solve_method == "diag"
    x = solver.solve(rhs)
    x_ref = ob.thomas_direct([(np.zeros((m, m)) if i == 0 else S[i][0],
                               np.zeros((m, m)) if i == nb - 1 else S[i][-1]) for i in range(nb)],
                             Tb, rhs)
    assert rel(x, x_ref) < 1e-10


@pytest.mark.parametrize("nb", [3, 6])
def test_multislab_cyclic_functions(T, nb):
    from sslablu_amd.multislab import omsdirectsolve as msd
    m = 20
    rng = np.random.default_rng(nb)
    S = [[0.3 * rng.standard_normal((m, m)) / np.sqrt(m),
          0.3 * rng.standard_normal((m, m)) / np.sqrt(m)] for _ in range(nb)]
    rhs = rng.standard_normal((nb * m, 2))
    dense = ob.assemble_dense([tuple(s) for s in S], [np.eye(m)] * nb, cyclic=True)
    x_ref = np.linalg.solve(dense, rhs)
    factors, smw = msd.build_block_cyclic_tridiagonal_solver(None, S, None, nb * m, m)
    x = msd.block_cyclic_tridiagonal_solve(None, factors, smw, rhs)
    assert rel(x, x_ref) < 1e-10
    # the non-periodic pair of functions on the same blocks (corners dropped)
    Sn = [[S[0][-1]]] + [S[i] for i in range(1, nb - 1)] + [[S[-1][0]]]
    xn = msd.block_tridiagonal_solve(None, msd.build_block_tridiagonal_solver(Sn), rhs)
    dn = ob.assemble_dense([(np.zeros((m, m)) if i == 0 else S[i][0],
                             np.zeros((m, m)) if i == nb - 1 else S[i][-1]) for i in range(nb)],
                           [np.eye(m)] * nb)
    assert rel(xn, np.linalg.solve(dn, rhs)) < 1e-10


def test_old_hbs_thomas_api_vs_oracle(T):
    """multislab/omsdirectsolveHBS.py (the older Thomas API with tree, quad and rank
    explicit) on config-1 blocks compressed accurately (rank 40, leaves of 40): against the
    oracle's dense FP64 solve of its own oms S-system and the exact solution."""
    from sslablu_amd.geometry.geom_2D import square
    from sslablu_amd.matAssembly.matAssembler import rkHMatAssembler
    from sslablu_amd.multislab import omsdirectsolveHBS as old_hbs
    from sslablu_amd.multislab.oms import oms
    from sslablu_amd.pdo.pdo import PDO2d, const
    from sslablu_amd.solver.solver import solverOptions
    N, q, ny = 8, 12, 16
    dSlabs, conn, H = square.dSlabs(N)
    opts = solverOptions("hps", [q, q], a=[H / 2, 1 / (2 * ny)])
    pdo = PDO2d(c11=const(1), c22=const(1))
    OMS = oms(dSlabs, pdo, square.gb_np, opts, conn)
    S_rk_list, rhs_list, Ntot, nc = OMS.construct_Stot_helper(exp_sin, rkHMatAssembler(40, 40))
    _, rhstot = OMS.construct_Stot_and_rhstot_linearOperator(S_rk_list, rhs_list, Ntot, nc)
    Tf = old_hbs.build_block_tridiagonal_solver(S_rk_list, S_rk_list[0][0].tree,
                                                S_rk_list[0][0].quad, 40)
    x = old_hbs.block_tridiagonal_solve(OMS, Tf, rhstot)
    ref = oh.oms_stot(pdo, N, (H, 1 / ny), q, bc=exp_sin)
    u = np.linalg.solve(oh.stot_dense(ref["S_rk_list"]), np.concatenate(ref["rhs_list"]))
    assert rel(x, u) < 1e-8
    assert np.max(np.abs(x - exp_sin(np.concatenate(ref["Ic_pts"])))) < 1e-7
