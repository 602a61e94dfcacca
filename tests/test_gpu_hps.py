"""GPU parity of the HPS S-assembly and the end-to-end oms solve.

* small grids: S blocks and right-hand sides vs the sparse oracle (oracle/hps.py),
  tolerance 1e-10 (relative, FP64);
* config 1 (Laplace, 8 slabs, p = 10, 16 leaves per interface -> m = 160): interface
  solution vs the exact u = exp(x) sin(y) and vs the oracle;
* config 2 size (32 slabs, p = 16, 32 leaves per interface -> m = 512): constant-kappa
  Helmholtz plane waves (a closed form) and the variable-kappa operator's residual.
"""
import numpy as np
import pytest

from oracle import blocktri as ob
from oracle import hps as oh
from tests.test_oracle_hps import LAPLACE, PDO2d, exp_sin, helmholtz, ones

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


def build(pdo, N, nx, ny, q):
    from sslablu_amd.hps import HPSSlabOperator, SlabGrid2D
    return HPSSlabOperator(pdo, SlabGrid2D(N, nx, ny, q)).build()


@pytest.mark.parametrize("N,nx,ny,q,pdo", [(3, 1, 2, 6, "laplace"), (4, 2, 2, 5, "helmholtz"),
                                           (5, 2, 4, 8, "helmholtz"), (4, 1, 8, 12, "laplace")])
def test_s_blocks_vs_oracle(T, N, nx, ny, q, pdo):
    pd = LAPLACE if pdo == "laplace" else helmholtz(6.0)
    op = build(pd, N, nx, ny, q)
    g = op.g
    ref = oh.oms_stot(pd, N, (g.hx, g.hy), q)
    got = op.S_rk_list()
    for i, blocks in enumerate(ref["S_rk_list"]):
        for Bg, Br in zip(got[i], blocks):
            Bg = Bg.cpu().numpy()
            assert np.linalg.norm(Bg - Br) / np.linalg.norm(Br) < 1e-10
    bc = lambda P: np.cos(P[:, 0] + 2 * P[:, 1]) + P[:, 0] ** 2
    F = T.tensor(bc(g.gb_points())[:, None], device=op.device)
    rhs = op.rhs_from_boundary(F)[:, 0].cpu().numpy()
    ref_rhs = np.concatenate([-R @ bc(X) for R, X in zip(ref["R_list"], ref["XXgb_list"])])
    assert np.linalg.norm(rhs - ref_rhs) / np.linalg.norm(ref_rhs) < 1e-10


def solve_oms(op, F, method="rb"):
    import torch
    from sslablu_amd.blocktri import BlockTriFactor
    g = op.g
    eye = torch.eye(g.m, dtype=torch.float64, device=op.device).expand(g.nb, g.m, g.m)
    fac = BlockTriFactor(method).factorize(eye, op.SiM, op.SiP)
    return fac.solve(op.rhs_from_boundary(F))


def test_config1_exact_solution(T):
    """Config 1: -Lap u = 0, u = exp(x) sin(y), 8 slabs, p = 10, 16 leaves / interface."""
    op = build(LAPLACE, 8, 2, 16, 12)
    g = op.g
    assert g.m == 160 and g.nb == 7
    F = T.tensor(exp_sin(g.gb_points()), device=op.device)
    u = solve_oms(op, F).cpu().numpy()[:, 0]
    exact = exp_sin(g.interface_points())
    assert np.max(np.abs(u - exact)) < 1e-10
    # and against the oracle's full pipeline (sparse slab solves + dense S solve)
    ref = oh.oms_stot(LAPLACE, 8, (g.hx, g.hy), 12, bc=exp_sin)
    u_ref = np.linalg.solve(oh.stot_dense(ref["S_rk_list"]), np.concatenate(ref["rhs_list"]))
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-10


@pytest.mark.parametrize("theta", [0.3, 1.1])
def test_config2_size_plane_wave(T, theta):
    """Config-2 sized grid (32 slabs, p = 16, m = 512) with constant kappa = 20: the plane
    wave u = sin(kappa (x cos t + y sin t)) solves -Lap u - kappa^2 u = 0 exactly."""
    kappa = 20.0
    pd = PDO2d(ones, ones, c=lambda P: -kappa ** 2 * np.ones(P.shape[0]))
    op = build(pd, 32, 1, 32, 18)
    g = op.g
    assert g.m == 512 and g.nb == 31
    u_ex = lambda P: np.sin(kappa * (P[:, 0] * np.cos(theta) + P[:, 1] * np.sin(theta)))
    F = T.tensor(u_ex(g.gb_points()), device=op.device)
    u = solve_oms(op, F).cpu().numpy()[:, 0]
    assert np.max(np.abs(u - u_ex(g.interface_points()))) < 1e-9


def test_config2_variable_helmholtz_residual(T):
    """Config 2 operator (variable kappa): RB and Thomas agree and the S-system residual
    is at FP64 level for 64 N(0,1) boundary-data columns."""
    import torch
    op = build(helmholtz(20.0), 32, 1, 32, 18)
    g = op.g
    rng = np.random.default_rng(0)
    F = torch.tensor(rng.standard_normal((g.gb_points().shape[0], 64)), device=op.device)
    rhs = op.rhs_from_boundary(F)
    x_rb = solve_oms(op, F, "rb")
    x_th = solve_oms(op, F, "thomas")
    assert (torch.linalg.norm(x_rb - x_th) / torch.linalg.norm(x_th)).item() < 1e-10
    # residual of u + S_l u_l + S_r u_r = rhs
    m, nb = g.m, g.nb
    X = x_rb.view(nb, m, -1)
    res = X.clone()
    res[1:] += torch.bmm(op.SiM[1:], X[:-1])
    res[:-1] += torch.bmm(op.SiP[:-1], X[1:])
    res = res.reshape(nb * m, -1) - rhs
    assert (torch.linalg.norm(res) / torch.linalg.norm(rhs)).item() < 1e-12
