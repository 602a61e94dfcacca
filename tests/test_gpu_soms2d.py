"""GPU: the 2-D SOMS system (sslablu_amd/multislab/soms.py, mirror of test/rkprop/SOMS.py)
against the oracle (oracle/soms2d.py): the dense Stot, the unknowns and index sets, the
block-sparse apply, the dense-LU and GMRES solves, exactness on a harmonic cubic and a
Helmholtz plane wave."""
import numpy as np
import pytest

from oracle import soms2d as so
from tests.test_oracle_soms2d import harmonic, plane

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


@pytest.mark.parametrize("px,py,nbx,nby,kh", [(8, 8, 3, 2, 0.0), (10, 6, 2, 4, 5.0),
                                              (6, 10, 1, 3, 0.0), (12, 12, 4, 4, 12.0)])
def test_system_matches_oracle(T, px, py, nbx, nby, kh):
    from sslablu_amd.multislab.soms import SOMS2D, SOMS_solver
    print(f"[soms] px={px} py={py} nbx={nbx} nby={nby} kh={kh}", flush=True)
    S, XY, Ii, Ib = SOMS_solver(px, py, nbx, nby, 1.5, 1.0, kh)
    So, XYo, Iio, Ibo = so.soms_system(px, py, nbx, nby, 1.5, 1.0, kh)
    assert np.array_equal(Ii, Iio) and np.array_equal(Ib, Ibo) and np.allclose(XY, XYo, atol=1e-15)
    Sd = S.cpu().numpy()
    assert np.linalg.norm(Sd - So) / np.linalg.norm(So) < 1e-12
    op = SOMS2D(px, py, nbx, nby, 1.5, 1.0, kh).build()
    X = T.tensor(np.random.default_rng(0).standard_normal((op.n, 3)), device="cuda")
    assert (T.linalg.norm(op.apply(X) - S @ X) / T.linalg.norm(S @ X)).item() < 1e-13
    ub = np.random.default_rng(1).standard_normal((len(Ib), 2))
    ref = so.soms_solve(So, Iio, Ibo, ub)
    got = op.solve(ub)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-11
    gm = op.solve(ub, method="gmres", tol=1e-13)
    assert np.linalg.norm(gm - ref) / np.linalg.norm(ref) < 1e-10


def test_exact_solutions(T):
    from sslablu_amd.multislab.soms import SOMS2D
    op = SOMS2D(8, 8, 3, 2, 1.5, 1.0, 0.0).build()
    XY, Ii, Ib = op.index_sets()
    u = harmonic(XY)
    assert np.max(np.abs(op.solve(u[Ib]) - u[Ii])) < 1e-10
    op = SOMS2D(16, 16, 3, 3, 1.0, 1.0, 12.0).build()
    XY, Ii, Ib = op.index_sets()
    u = plane(XY, 12.0)
    assert np.max(np.abs(op.solve(T.tensor(u[Ib], device="cuda")).cpu().numpy() - u[Ii])) < 1e-10


def test_noms_mirror_matches_oms(T):
    """multislab/noms.py mirror: slab-by-slab construction through solverWrapper's
    solver_ii / Aib gives the S-system of the fused oms assembly and the exact solution."""
    from sslablu_amd.geometry.geom_2D import square
    from sslablu_amd.matAssembly.matAssembler import denseMatAssembler
    from sslablu_amd.multislab import noms, oms
    from sslablu_amd.pdo.pdo import PDO2d, const
    from sslablu_amd.solver.solver import solverOptions

    def exp_sin(P):
        return np.exp(P[:, 0]) * np.sin(P[:, 1])

    N, q, ny = 5, 10, 4
    dSlabs, conn, H = square.dSlabs(N)
    opts = solverOptions("hps", [q, q], a=[H / 2, 1 / (2 * ny)])
    pdo = PDO2d(c11=const(1), c22=const(1))
    A, rhs = noms.oms(dSlabs, pdo, square.gb_np, opts, conn).construct_Stot_and_rhstot(
        exp_sin, denseMatAssembler())
    A0, rhs0 = oms.oms(dSlabs, pdo, square.gb_np, opts, conn).construct_Stot_and_rhstot(
        exp_sin, denseMatAssembler())
    assert np.linalg.norm(rhs - rhs0) / np.linalg.norm(rhs0) < 1e-12
    n = A.shape[0]
    D, D0 = A @ np.eye(n), A0 @ np.eye(n)
    assert np.linalg.norm(D - D0) / np.linalg.norm(D0) < 1e-12
    u = np.linalg.solve(D, rhs)
    ref = oms.oms(dSlabs, pdo, square.gb_np, opts, conn)
    ref.construct_Stot_and_rhstot(exp_sin, denseMatAssembler())
    assert np.max(np.abs(u - exp_sin(ref.op.g.interface_points()))) < 1e-8
    DT = A.rmatmat(np.eye(n))
    assert np.linalg.norm(DT - D.T) / np.linalg.norm(D) < 1e-12
