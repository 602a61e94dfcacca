"""GPU parity of the 2-D spectral double slabs with variable coefficients
(sslablu_amd/spectral2d.py, sslab_spectral2d_slab_maps) against the oracle
(oracle/spectral2d.py): S blocks and right-hand sides (1e-10), and the oms mirror with
solverOptions('spectral', ord) end to end -- 2-D (variable coefficients) and 3-D
(constant coefficients, fast diagonalization) -- against the oracles' dense solves and
exact solutions."""
import numpy as np
import pytest

from oracle import spectral2d as os2

pytestmark = pytest.mark.gpu


class PDO2:
    def __init__(self, c=None, c1=None, c2=None, c11=None):
        one = lambda P: np.ones(P.shape[0])  # noqa: E731
        self.c11 = c11 or one
        self.c22 = one
        self.c1, self.c2, self.c = c1, c2, c
        self.c12 = None


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def poly(P):
    return P[:, 0] ** 2 - P[:, 1] ** 2 + 3 * P[:, 0] * P[:, 1] + P[:, 0]


@pytest.mark.parametrize("N,ordx,ordy,kind", [(5, 6, 12, "laplace"), (4, 8, 20, "variable"),
                                              (7, 4, 9, "helmholtz")])
def test_blocks_and_rhs_vs_oracle(gpu, N, ordx, ordy, kind):
    import torch
    from sslablu_amd.spectral2d import SpectralSlabs2D
    pdo = {"laplace": PDO2(),
           "variable": PDO2(c=lambda P: -4.0 * (1 + P[:, 0] * P[:, 1]),
                            c1=lambda P: 0.5 + P[:, 1], c2=lambda P: np.cos(P[:, 0]),
                            c11=lambda P: 1.0 + 0.25 * np.sin(P[:, 0] + P[:, 1])),
           "helmholtz": PDO2(c=lambda P: -25.0 * (1 - 0.5 * np.exp(-((P - 0.5) ** 2).sum(1) / 0.05)))}[kind]
    op = SpectralSlabs2D(pdo, N, ordx, ordy).build()
    assert op.info_bad == 0
    ref = os2.oms_stot(pdo, N, ordx, ordy, bc=poly)
    S = ref["S_rk_list"]
    nb = N - 1
    for i in range(nb):
        if i > 0:
            assert rel(op.SiM[i], S[i][0]) < 1e-10
        if i < nb - 1:
            assert rel(op.SiP[i], S[i][-1]) < 1e-10
    assert rel(op.rhs_from_function(poly)[:, 0], np.concatenate(ref["rhs_list"])) < 1e-10
    assert np.allclose(op.interface_points(), np.concatenate(ref["Ic_pts"]), atol=1e-15)


def test_oms_spectral_2d_and_3d(gpu):
    from sslablu_amd.direct_solve.omsdirectsolve import RedBlackSolver
    from sslablu_amd.geometry.geom_2D import square
    from sslablu_amd.geometry.geom_3D import cube
    from sslablu_amd.matAssembly.matAssembler import denseMatAssembler
    from sslablu_amd.multislab.oms import oms
    from sslablu_amd.pdo import pdo
    from sslablu_amd.solver.solver import solverOptions
    # 2-D, variable coefficients: the oms S-system equals the oracle's
    var = PDO2(c=lambda P: -3.0 * (1 + P[:, 0]), c11=lambda P: 1.0 + 0.2 * P[:, 1])
    dSlabs, conn, H = square.dSlabs(6)
    OMS = oms(dSlabs, var, square.gb_np, solverOptions("spectral", [8, 16]), conn)
    S_rk, rhs_list, Ntot, nc = OMS.construct_Stot_helper(poly, denseMatAssembler())
    solver = RedBlackSolver(nc)
    solver.factorize(S_rk, [np.eye(nc)] * len(S_rk))
    x = solver.solve(np.concatenate(rhs_list))
    ref = os2.oms_stot(var, 6, 8, 16, bc=poly)
    from oracle.hps import stot_dense
    u = np.linalg.solve(stot_dense(ref["S_rk_list"]), np.concatenate(ref["rhs_list"]))
    assert rel(x, u) < 1e-10
    # Laplace: exact on a harmonic polynomial
    lap = pdo.PDO2d(c11=pdo.const(1), c22=pdo.const(1))
    OMS = oms(dSlabs, lap, square.gb_np, solverOptions("spectral", [8, 16]), conn)
    S_rk, rhs_list, Ntot, nc = OMS.construct_Stot_helper(poly, denseMatAssembler())
    solver.factorize(S_rk, [np.eye(nc)] * len(S_rk))
    x = solver.solve(np.concatenate(rhs_list))
    pts = OMS.discretize().interface_points()
    assert np.max(np.abs(x - poly(pts))) < 1e-11
    # 3-D, constant coefficients through the same options
    dS3, conn3, H3 = cube.dSlabs(4)
    lap3 = pdo.PDO3d(c11=pdo.const(1), c22=pdo.const(1), c33=pdo.const(1))
    O3 = oms(dS3, lap3, cube.gb_np, solverOptions("spectral", [6, 9, 9]), conn3)
    S3, r3, N3, nc3 = O3.construct_Stot_helper(
        lambda P: P[:, 0] ** 2 - P[:, 2] ** 2 + P[:, 1], denseMatAssembler())
    s3 = RedBlackSolver(nc3)
    s3.factorize(S3, [np.eye(nc3)] * len(S3))
    x3 = s3.solve(np.concatenate(r3))
    P3 = O3.discretize().interface_points()
    assert np.max(np.abs(x3 - (P3[:, 0] ** 2 - P3[:, 2] ** 2 + P3[:, 1]))) < 1e-10
