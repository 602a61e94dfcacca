"""GPU: x-periodic slab decompositions (the ring connectivity of
geometry/geom_3D/squareTorus.py:268-277) in 2-D and 3-D against the periodic oracles
(oracle/hps.py, oracle/hps3d.py with periodic=True, pinned by exact domain
decomposition against a global solve with the end lines / faces identified): S blocks and
right-hand sides (1e-10), the cyclic S-system solved on the GPU (Woodbury-corrected
Thomas and red-black factors, the multislab cyclic functions, RedBlackSolver(cyclic))
against the oracle's dense cyclic solve (1e-10), and the oms mirror with periodic
connectivity."""
import numpy as np
import pytest

from oracle import hps as oh
from oracle import hps3d as o3

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


class PDO2:
    c11 = c22 = staticmethod(lambda P: np.ones(P.shape[0]))
    c1 = c2 = c12 = None

    def __init__(self, c=None):
        self.c = c


class PDO3:
    c11 = c22 = c33 = staticmethod(lambda P: np.ones(P.shape[0]))
    c1 = c2 = c3 = None

    def __init__(self, c=None):
        self.c = c


def u2(P):
    return np.sin(2 * np.pi * P[:, 0]) * np.exp(2 * np.pi * (P[:, 1] - 1))


def u3(P):
    return np.sin(2 * np.pi * P[:, 0]) * np.sin(np.pi * P[:, 2]) * np.exp(np.sqrt(5) * np.pi * (P[:, 1] - 1))


@pytest.mark.parametrize("N,c", [(4, None), (5, lambda P: -30.0 * (1 + 0.3 * np.sin(2 * np.pi * P[:, 0])))])
def test_periodic_2d(gpu, N, c):
    import torch
    from sslablu_amd.blocktri import BlockTriFactor
    from sslablu_amd.hps import HPSSlabOperator, SlabGrid2D
    pdo = PDO2(c)
    q, nx, ny = 8, 2, 4
    g = SlabGrid2D(N, nx, ny, q, periodic=True)
    op = HPSSlabOperator(pdo, g).build()
    ref = oh.oms_stot(pdo, N, (g.hx, g.hy), q, bc=u2, periodic=True)
    for i in range(N):
        assert rel(op.SiM[i], ref["S_rk_list"][i][0]) < 1e-10
        assert rel(op.SiP[i], ref["S_rk_list"][i][1]) < 1e-10
    assert np.allclose(g.interface_points(), np.concatenate(ref["Ic_pts"]), atol=1e-15)
    B = op.rhs_from_boundary(torch.tensor(u2(g.gb_points()), device="cuda"))
    assert rel(B[:, 0], np.concatenate(ref["rhs_list"])) < 1e-10
    u = np.linalg.solve(oh.stot_dense_cyclic(ref["S_rk_list"]), np.concatenate(ref["rhs_list"]))
    for method in ("thomas", "rb"):
        fac = BlockTriFactor(method).factorize(None if method == "rb" else None, op.SiM, op.SiP,
                                                E=op.SiM[0], F=op.SiP[N - 1])
        assert rel(fac.solve(B)[:, 0], u) < 1e-10


def test_periodic_3d_and_oms_mirror(gpu):
    import torch
    from sslablu_amd.direct_solve.omsdirectsolve import BlockTridiagonalSolver, RedBlackSolver
    from sslablu_amd.geometry.geom_3D import squareTorus
    from sslablu_amd.hps3d import HPSSlabOperator3D, SlabGrid3D
    from sslablu_amd.matAssembly.matAssembler import denseMatAssembler
    from sslablu_amd.multislab import omsdirectsolve as msd
    from sslablu_amd.multislab.oms import oms
    from sslablu_amd.solver.solver import solverOptions
    N, q = 4, 5
    g = SlabGrid3D(N, 1, 2, 2, q, periodic=True)
    op = HPSSlabOperator3D(PDO3(), g).build()
    ref = o3.oms_stot(PDO3(), N, (g.hx, g.hy, g.hz), q, bc=u3, periodic=True)
    SiM, SiP = op.dense_blocks()
    for i in range(N):
        assert rel(SiM[i], ref["S_rk_list"][i][0]) < 1e-10
        assert rel(SiP[i], ref["S_rk_list"][i][1]) < 1e-10
    B = op.rhs_from_boundary(torch.tensor(u3(g.gb_points()), device="cuda"))
    assert rel(B[:, 0], np.concatenate(ref["rhs_list"])) < 1e-10
    u = np.linalg.solve(oh.stot_dense_cyclic(ref["S_rk_list"]), np.concatenate(ref["rhs_list"]))
    A = op.tridiag_apply()
    assert rel(A(torch.tensor(u[:, None], device="cuda"))[:, 0], np.concatenate(ref["rhs_list"])) < 1e-10
    # the oms mirror with squareTorus's ring connectivity, then the reference's periodic
    # solve routes: multislab cyclic Thomas + SMW, and the direct_solve cyclic solvers
    dS, conn, H = squareTorus.dSlabs(N)
    opts = solverOptions("hps", [q] * 3, [H / 2, 1 / 4, 1 / 4])
    OMS = oms(dS, PDO3(), squareTorus.gb_np, opts, conn)
    S_rk, rhs_list, Ntot, nc = OMS.construct_Stot_helper(u3, denseMatAssembler())
    assert Ntot == N * nc and all(len(b) == 2 for b in S_rk)
    rhs = np.concatenate(rhs_list)
    assert rel(rhs, np.concatenate(ref["rhs_list"])) < 1e-10
    T, smw = msd.build_block_cyclic_tridiagonal_solver(OMS, S_rk, rhs_list, Ntot, nc)
    assert rel(msd.block_cyclic_tridiagonal_solve(OMS, T, smw, rhs), u) < 1e-10
    Sd = [[np.asarray(b) for b in s] for s in S_rk]
    for cls in (BlockTridiagonalSolver, RedBlackSolver):
        slv = cls(nc, cyclic=True)
        slv.factorize(Sd, [np.eye(nc)] * N)
        assert rel(slv.solve(rhs), u) < 1e-10
