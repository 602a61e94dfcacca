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


@pytest.mark.parametrize("N,nx,ny,q,r", [(3, 1, 2, 6, 3), (4, 2, 2, 5, 1), (5, 2, 4, 8, 7),
                                         (6, 1, 4, 6, 5), (9, 2, 2, 6, 4), (10, 1, 2, 5, 2)])
def test_rhs_rb0_matches_two_step(T, N, nx, ny, q, r):
    """The fused right-hand side + first red-black level (sslab_oms_rhs_rb0) equals the
    two-step form (rhs, then v_i - SiM_i v_{i-1} - SiP_i v_{i+1} on even i; odd rows
    rhs_j into X, X's even rows zero), and the solve from it (SSLAB_RB_LEVEL0_REDUCED)
    equals the plain solve, for even and odd interface counts nb = N - 1 (the right
    boundary then feeds an even or odd row).  The buffers start as NaN so every entry
    the entry state promises is shown to be written."""
    from sslablu_amd.blocktri import BlockTriFactor
    op = build(helmholtz(6.0), N, nx, ny, q)
    g = op.g
    m, nb = g.m, g.nb
    rng = np.random.default_rng(N * 10 + r)
    F = T.tensor(rng.standard_normal((g.gb_points().shape[0], r)), device=op.device)
    rhs = op.rhs_from_boundary(F)
    V = T.full((nb * m, r), float("nan"), dtype=T.float64, device=op.device)
    X = T.full_like(V, float("nan"))
    op.rhs_rb0_from_boundary(F, V, X)
    R = rhs.view(nb, m, r)
    want = R.clone()
    want[2::2] -= T.bmm(op.SiM[2::2], R[1:nb - 1:2])
    nO = nb // 2
    want[0:2 * nO:2] -= T.bmm(op.SiP[0:2 * nO:2], R[1::2])
    Vb, Xb = V.view(nb, m, r), X.view(nb, m, r)
    scale = T.linalg.norm(want).item()
    assert T.linalg.norm(Vb[0::2] - want[0::2]).item() < 1e-12 * scale
    assert T.linalg.norm(Xb[1::2] - R[1::2]).item() < 1e-13 * scale
    assert T.count_nonzero(Xb[0::2]).item() == 0
    fac = BlockTriFactor("rb", op.device).factorize(None, op.SiM, op.SiP)
    X0 = fac.solve(rhs)
    fac.solve_(V, X, level0_reduced=True)
    assert (T.linalg.norm(X - X0) / T.linalg.norm(X0)).item() < 1e-12


def test_rhs_rb0_exact_solution_and_errors(T):
    """Config 1 (Laplace, u = exp(x) sin(y)) through the fused path reproduces the exact
    interface values; the fused path rejects factors that are not red-black with T = I,
    and zero right-hand-side columns are a no-op."""
    import torch
    from sslablu_amd.blocktri import BlockTriFactor
    op = build(LAPLACE, 8, 2, 16, 12)
    g = op.g
    F = T.tensor(exp_sin(g.gb_points()), device=op.device)
    V, X = op.rhs_rb0_from_boundary(F)
    fac = BlockTriFactor("rb", op.device).factorize(None, op.SiM, op.SiP)
    fac.solve_(V, X, level0_reduced=True)
    assert np.max(np.abs(X.cpu().numpy()[:, 0] - exp_sin(g.interface_points()))) < 1e-10
    eye = torch.eye(g.m, dtype=torch.float64, device=op.device).expand(g.nb, g.m, g.m)
    fac_t = BlockTriFactor("rb", op.device).factorize(eye, op.SiM, op.SiP)
    with pytest.raises(ValueError):
        fac_t.solve_(*op.rhs_rb0_from_boundary(F), level0_reduced=True)
    V0, X0 = op.rhs_rb0_from_boundary(F[:, None][:, :0])
    assert V0.shape == (g.nb * g.m, 0) and X0.shape == (g.nb * g.m, 0)


@pytest.mark.parametrize("N,nx,ny,q,r", [(2, 1, 2, 6, 2), (3, 1, 2, 6, 3), (4, 2, 2, 5, 1),
                                         (5, 2, 4, 8, 7), (9, 2, 2, 6, 4), (10, 1, 2, 5, 2),
                                         (17, 1, 4, 6, 3)])
def test_boundary_fold_matches_plain(T, N, nx, ny, q, r):
    """The folded forward sweep (sslab_oms_fold_*): its entry state equals the
    forward-only red-black phase applied to the formed right-hand sides, and the folded
    solve equals the plain rhs + solve, for one to sixteen interfaces (every parity of
    the boundary nodes and of the coarsest level)."""
    from sslablu_amd.blocktri import BlockTriFactor
    op = build(helmholtz(6.0), N, nx, ny, q)
    g = op.g
    rng = np.random.default_rng(N * 31 + r)
    F = T.tensor(rng.standard_normal((g.gb_points().shape[0], r)), device=op.device)
    fac = BlockTriFactor("rb", op.device).factorize(None, op.SiM, op.SiP)
    fold = op.fold(fac)
    assert fold.groups >= 1 and fold.nbytes > 0
    rhs = op.rhs_from_boundary(F)
    E0 = T.empty_like(rhs)
    fac.solve_(rhs.clone(), E0, phase="forward")
    E1 = fold.apply_(F, T.full_like(rhs, float("nan")))
    assert (T.linalg.norm(E1 - E0) / T.linalg.norm(E0)).item() < 1e-12
    X0 = fac.solve(rhs)
    X1 = fold.solve(F)
    assert (T.linalg.norm(X1 - X0) / T.linalg.norm(X0)).item() < 1e-12


def test_boundary_fold_exact_and_config2(T):
    """Config 1 (u = exp(x) sin(y)) and the config-2 sized plane wave through the folded
    solve reproduce the exact interface values."""
    from sslablu_amd.blocktri import BlockTriFactor
    op = build(LAPLACE, 8, 2, 16, 12)
    g = op.g
    fac = BlockTriFactor("rb", op.device).factorize(None, op.SiM, op.SiP)
    u = op.fold(fac).solve(T.tensor(exp_sin(g.gb_points()), device=op.device))
    assert np.max(np.abs(u.cpu().numpy()[:, 0] - exp_sin(g.interface_points()))) < 1e-10
    kappa, theta = 20.0, 0.7
    pd = PDO2d(ones, ones, c=lambda P: -kappa ** 2 * np.ones(P.shape[0]))
    op = build(pd, 32, 1, 32, 18)
    g = op.g
    u_ex = lambda P: np.sin(kappa * (P[:, 0] * np.cos(theta) + P[:, 1] * np.sin(theta)))
    fac = BlockTriFactor("rb", op.device).factorize(None, op.SiM, op.SiP)
    u = op.fold(fac).solve(T.tensor(u_ex(g.gb_points()), device=op.device))
    assert np.max(np.abs(u.cpu().numpy()[:, 0] - u_ex(g.interface_points()))) < 1e-9


def _match(P, Q):
    """Index of every row of P in Q (both point lists; coordinates agree to 1e-9)."""
    from scipy.spatial import cKDTree
    d, idx = cKDTree(Q).query(P)
    assert np.max(d) < 1e-9
    return idx


def test_solve_dir_exact_and_global(T):
    """Interior reconstruction (the reference's solve_dir by the HPS downward pass):
    config 1's u = exp(x) sin(y) at every leaf-interior node; the variable-coefficient
    Helmholtz operator against the oracle's brute-force global HPS solve at the same
    nodes (exact domain decomposition); several columns at once equal one at a time;
    errors when the operators were not kept or a body load is given."""
    from sslablu_amd.blocktri import BlockTriFactor
    from sslablu_amd.hps import HPSSlabOperator, SlabGrid2D
    op = HPSSlabOperator(LAPLACE, SlabGrid2D(8, 2, 16, 12), keep_solution_ops=True).build()
    g = op.g
    F = T.tensor(exp_sin(g.gb_points()), device=op.device)
    fac = BlockTriFactor("rb", op.device).factorize(None, op.SiM, op.SiP)
    x = op.fold(fac).solve(F)
    u = op.solve_dir(x, F).cpu().numpy()[:, 0]
    assert u.shape[0] == g.interior_points().shape[0]
    assert np.max(np.abs(u - exp_sin(g.interior_points()))) < 1e-10

    pd = helmholtz(6.0)
    op = HPSSlabOperator(pd, SlabGrid2D(4, 2, 2, 8), keep_solution_ops=True).build()
    g = op.g
    bc = lambda P: np.cos(P[:, 0] + 2 * P[:, 1]) + P[:, 0] ** 2
    F1 = T.tensor(bc(g.gb_points()), device=op.device)
    fac = BlockTriFactor("rb", op.device).factorize(None, op.SiM, op.SiP)
    u = op.solve_dir(op.fold(fac).solve(F1)[:, 0], F1).cpu().numpy()
    assert u.ndim == 1
    ref, XX = oh.global_solution(pd, 4, (g.hx, g.hy), 8, bc)
    u_ref = ref[_match(g.interior_points(), XX)]
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-10
    rng = np.random.default_rng(3)
    F3 = T.tensor(rng.standard_normal((g.gb_points().shape[0], 3)), device=op.device)
    x3 = fac.solve(op.rhs_from_boundary(F3))
    U3 = op.solve_dir(x3, F3)
    for k in range(3):
        uk = op.solve_dir(x3[:, k], F3[:, k])
        assert (T.linalg.norm(U3[:, k] - uk) / T.linalg.norm(uk)).item() < 1e-13
    with pytest.raises(NotImplementedError):
        op.solve_dir(x3, F3, ff_body=np.zeros(3))
    plain = HPSSlabOperator(pd, SlabGrid2D(4, 2, 2, 8)).build()
    with pytest.raises(ValueError):
        plain.solve_dir(x3, F3)
