"""GPU: a driver shaped like the reference's 3-D cube test (test/3D/cubeTest.py:99-186)
runs on the mirrors: cube.dSlabs, solverOptions('hpsalt', [p_disc]*3, a), oms with the
rkHBS assembler, ThomasSolverHBS and RedBlackSolverHBS (5 interfaces: not a power of
two), oms_lu for the exact S-system operator, and scipy GMRES preconditioned by the HBS
solvers.  Checked against the 3-D oracle's dense S-system solution (1e-9) and the exact
harmonic solution at the interface nodes (the discretization error of q = 7 leaves,
1e-4)."""
import numpy as np
import pytest
import scipy
from scipy.sparse.linalg import LinearOperator, gmres

from oracle import hps3d as o3

pytestmark = pytest.mark.gpu


def u_exact(P):
    return np.exp(np.sqrt(2) * np.pi * P[:, 0]) * np.sin(np.pi * P[:, 1]) * np.sin(np.pi * P[:, 2])


def test_cube_driver(gpu):
    from sslablu_amd.direct_solve import omsdirectsolveHBS as omsdirectHBS
    from sslablu_amd.geometry.geom_3D import cube
    from sslablu_amd.matAssembly import matAssembler as mA
    from sslablu_amd.multislab import oms
    from sslablu_amd.pdo import pdo
    from sslablu_amd.solver import solver as solverWrap
    N, p_disc = 6, 7
    dSlabs, connectivity, H = cube.dSlabs(N)
    one = pdo.const(1)
    Lap = pdo.PDO3d(c11=one, c22=one, c33=one)
    a = np.array([H / 2, 1 / 4, 1 / 4])                     # leaf half-sizes: 2 x 2 per face
    opts = solverWrap.solverOptions("hpsalt", [p_disc] * 3, a)
    gb = lambda P: cube.gb(P, jax_avail=False, torch_avail=True)  # noqa: E731
    OMS = oms.oms(dSlabs, Lap, gb, opts, connectivity)
    S_rk_list, rhs_list, Ntot, nc = OMS.construct_Stot_helper(u_exact, mA.rkHMatAssembler(25, 20))
    assert nc == 4 * (p_disc - 2) ** 2 and Ntot == (N - 1) * nc
    OMS_lu = oms.oms_lu(dSlabs, Lap, gb, opts, connectivity)
    S_lu, rhs_lu, _, _ = OMS_lu.construct_Stot_helper(u_exact)
    Stot, rhstot = OMS_lu.construct_Stot_and_rhstot_linearOperator(S_lu, rhs_lu, Ntot, nc)
    assert np.allclose(rhstot, np.concatenate(rhs_list))
    ref = o3.oms_stot(Lap, N, 2 * a, p_disc, bc=u_exact)
    u_ref = np.linalg.solve(o3.stot_dense(ref["S_rk_list"]), np.concatenate(ref["rhs_list"]))
    pts = np.concatenate(ref["Ic_pts"])
    thomas = omsdirectHBS.ThomasSolverHBS(nc, 20)
    rb = omsdirectHBS.RedBlackSolverHBS(nc, 20, S_rk_list[0][0].tree, S_rk_list[0][0].quad)
    thomas.factorize(S_rk_list)
    rb.factorize(S_rk_list)
    kw = {"rtol": 1e-12} if scipy.__version__ >= "1.14" else {"tol": 1e-12}
    for slv in (thomas, rb):
        M = LinearOperator((Ntot, Ntot), matvec=slv.solve, dtype=np.float64)
        x, info = gmres(Stot, rhstot, M=M, maxiter=60, restart=60, **kw)
        assert info == 0
        assert np.linalg.norm(Stot @ x - rhstot) / np.linalg.norm(rhstot) < 1e-10
        assert np.linalg.norm(x - u_ref) / np.linalg.norm(u_ref) < 1e-9
        assert np.max(np.abs(x - u_exact(pts))) / np.max(np.abs(u_exact(pts))) < 1e-4
