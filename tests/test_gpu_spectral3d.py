"""GPU: the 3-D spectral slab S-system (fast diagonalization on the device) against the
dense-assembly oracle, the exact harmonic solution, and its HBS-compressed solve.
Tolerances: S blocks / rhs vs oracle 1e-10 relative; interface solution vs the exact
u = exp(sqrt(2) pi x) sin(pi y) sin(pi z) 1e-8 (spectral discretization); HBS solve vs the
dense FP64 solve within 100x the compression error of the blocks."""
import numpy as np
import pytest

from oracle import spectral3d as o3

pytestmark = pytest.mark.gpu


def g_exp(P):
    return np.exp(np.sqrt(2) * np.pi * P[:, 0]) * np.sin(np.pi * P[:, 1]) * np.sin(np.pi * P[:, 2])


def g_two(P):
    return np.stack([g_exp(P), P[:, 0] ** 2 - P[:, 1] ** 2 + 3 * P[:, 2] * P[:, 0]], axis=1)


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else a
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("N,nx,ny,nz", [(4, 5, 8, 7), (3, 7, 6, 9), (2, 3, 5, 5)])
def test_blocks_and_rhs_vs_oracle(T, N, nx, ny, nz):
    from sslablu_amd.spectral3d import SpectralSlabs3D
    SiM, SiP, rhs, pts = o3.oms_spectral3d(N, nx, ny, nz, g_two)
    s = SpectralSlabs3D(N, nx, ny, nz).build()
    if N > 2:  # with one interface the oracle's couplings are all boundary (zero)
        assert rel(s.S_left, SiM[1]) < 1e-10
        assert rel(s.S_right, SiP[0]) < 1e-10
    assert rel(s.rhs_from_function(g_two), rhs) < 1e-10
    assert np.allclose(s.interface_points(), pts)


def test_exact_solution_dense(T):
    from sslablu_amd.blocktri import BlockTriFactor
    from sslablu_amd.spectral3d import SpectralSlabs3D
    s = SpectralSlabs3D(8, 9, 34, 34).build()
    SiM, SiP = s.blocks()
    fac = BlockTriFactor("rb").factorize(None, SiM, SiP)
    x = fac.solve(s.rhs_from_function(g_exp))[:, 0].cpu().numpy()
    ex = g_exp(s.interface_points())
    assert np.abs(x - ex).max() / np.abs(ex).max() < 1e-8


def test_hbs_compressed_thomas_solve(T):
    import torch
    from sslablu_amd.blocktri import BlockTriFactor
    from sslablu_amd.direct_solve.omsdirectsolveHBS import ThomasSolverHBS
    from sslablu_amd.matAssembly.HBS import HBSnew
    from sslablu_amd.spectral3d import FaceTree, SpectralSlabs3D
    s = SpectralSlabs3D(8, 9, 34, 34).build()
    tree = FaceTree(32, 32, 64)
    assert tree.nleaves == 16 and len(tree.perm_leaf) == 1024
    rk = 30
    H = {}
    for side in ("left", "right"):
        h = HBSnew.HBSMAT(s.S[side], tree, quad=False, samples=max(2 * rk, 64) + rk + 10)
        h.construct(rk)
        H[side] = h
    I = torch.eye(s.m, dtype=torch.float64, device="cuda")
    errs = [rel(H[sd].matmat(I), s.S[sd].cpu().numpy()) for sd in ("left", "right")]
    nb = s.nb
    S_rk = [[H["right"]]] + [[H["left"], H["right"]] for _ in range(nb - 2)] + [[H["left"]]]
    B = s.rhs_from_function(g_exp)
    solver = ThomasSolverHBS(s.m, rk)
    solver.factorize(S_rk)
    x_hbs = solver.solve(B)
    SiM, SiP = s.blocks()
    x_dense = BlockTriFactor("rb").factorize(None, SiM, SiP).solve(B)
    e = rel(x_hbs, x_dense.cpu().numpy())
    assert e < max(100 * max(errs), 1e-10), (errs, e)


def test_hbs_solve_graph_replay(T):
    """The HBS Thomas solve (ULV solves and HBS applies, thousands of small launches at
    the C3 size) captured once as a HIP graph: every replay, with new right-hand sides
    copied into the captured input and eager work in between, equals the eager solve."""
    import torch
    from sslablu_amd.direct_solve.omsdirectsolveHBS import ThomasSolverHBS
    from sslablu_amd.matAssembly.HBS import HBSnew
    from sslablu_amd.spectral3d import FaceTree, SpectralSlabs3D
    s = SpectralSlabs3D(6, 9, 34, 34).build()
    tree = FaceTree(32, 32, 64)
    rk = 20
    H = {}
    for side in ("left", "right"):
        h = HBSnew.HBSMAT(s.S[side], tree, quad=False, seed=2,
                          samples=max(2 * rk, 64) + rk + 10)
        h.construct(rk)
        H[side] = h
    nb = s.nb
    S_rk = [[H["right"]]] + [[H["left"], H["right"]] for _ in range(nb - 2)] + [[H["left"]]]
    solver = ThomasSolverHBS(s.m, rk, samples=max(2 * rk, 64) + rk + 10)
    solver.factorize(S_rk)
    from sslablu_amd.graphs import CapturedSolve
    gen = torch.Generator(device="cuda").manual_seed(5)
    cs = CapturedSolve(solver.solve, (nb * s.m, 4))
    for rep in range(3):
        B = torch.randn(nb * s.m, 4, dtype=torch.float64, device="cuda", generator=gen)
        X_out = cs(B)
        torch.cuda.synchronize()
        X_eager = torch.as_tensor(solver.solve(B), device="cuda")
        assert (torch.linalg.norm(X_out - X_eager) / torch.linalg.norm(X_eager)).item() < 1e-13
    with pytest.raises(ValueError):
        cs(B[:, :2])
