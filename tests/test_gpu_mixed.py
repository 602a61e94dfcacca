"""GPU parity of the mixed-precision path: FP32 MFMA GEMM, FP32 red-black solve as the
preconditioner, FP64 residual, iterative refinement.

Tolerances: FP64 residual bit-level against numpy (1e-13 relative); refined solutions
<= 1e-8 relative error vs the FP64 solve (the stated tolerance of the mixed form), and
<= 1e-10 after two refinement steps on the well-conditioned systems used here.
"""
import numpy as np
import pytest

from oracle import blocktri as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


def _system(nb, m, seed, identity=True):
    rng = np.random.default_rng(seed)
    Tb = [np.eye(m) if identity else np.eye(m) + 0.2 * rng.standard_normal((m, m)) / np.sqrt(m)
          for _ in range(nb)]
    S = [(0.4 * rng.standard_normal((m, m)) / np.sqrt(m) if i > 0 else np.zeros((m, m)),
          0.4 * rng.standard_normal((m, m)) / np.sqrt(m) if i < nb - 1 else np.zeros((m, m)))
         for i in range(nb)]
    return Tb, S


@pytest.mark.parametrize("identity", [True, False])
def test_residual(T, identity):
    from sslablu_amd.blocktri import BlockTriFactor
    nb, m, r = 5, 24, 3
    Tb, S = _system(nb, m, 1, identity)
    A = ob.assemble_dense(S, Tb)
    rng = np.random.default_rng(2)
    B = rng.standard_normal((nb * m, r))
    X = rng.standard_normal((nb * m, r))
    fac = BlockTriFactor("rb").factorize(np.stack(Tb), np.stack([s[0] for s in S]),
                                         np.stack([s[1] for s in S]))
    fac.enable_mixed(None if identity else np.stack(Tb), np.stack([s[0] for s in S]),
                     np.stack([s[1] for s in S]))
    R = fac.residual(T.tensor(B, device="cuda"), T.tensor(X, device="cuda")).cpu().numpy()
    ref = B - A @ X
    assert np.linalg.norm(R - ref) / np.linalg.norm(ref) < 1e-13


@pytest.mark.parametrize("nb,m,r", [(7, 64, 4), (31, 128, 16), (12, 96, 5)])
def test_mixed_refinement(T, nb, m, r):
    from sslablu_amd.blocktri import BlockTriFactor
    Tb, S = _system(nb, m, nb + m)
    SiM = np.stack([s[0] for s in S])
    SiP = np.stack([s[1] for s in S])
    fac = BlockTriFactor("rb").factorize(np.stack(Tb), SiM, SiP).enable_mixed(None, SiM, SiP)
    rng = np.random.default_rng(3)
    B = T.tensor(rng.standard_normal((nb * m, r)), device="cuda")
    x64 = fac.solve(B)
    errs = []
    for iters in (0, 1, 2):
        X = T.empty_like(B)
        fac.solve_mixed_(B.contiguous(), X, iters=iters)
        errs.append((T.linalg.norm(X - x64) / T.linalg.norm(x64)).item())
    assert 1e-9 < errs[0] < 1e-4          # FP32 preconditioner alone
    assert errs[1] < errs[0] * 1e-2       # refinement contracts
    assert errs[2] < 1e-10
    # and the FP64 solve itself matches the oracle
    x_ref = ob.thomas_direct(S, Tb, B.cpu().numpy())
    assert np.linalg.norm(x64.cpu().numpy() - x_ref) / np.linalg.norm(x_ref) < 1e-12


@pytest.mark.parametrize("identity", [True, False])
@pytest.mark.parametrize("nb,m,r", [(31, 128, 8), (9, 200, 3)])
def test_fp32_factorization_refined(T, identity, nb, m, r):
    """Factor in FP32 (FP32 LU + FP32 MFMA Schur updates), refine with FP64 residuals:
    <= 1e-8 vs the FP64 solve after one step, <= 1e-10 after two."""
    from sslablu_amd.blocktri import BlockTriFactor
    Tb, S = _system(nb, m, 7 + nb, identity)
    SiM = np.stack([s[0] for s in S])
    SiP = np.stack([s[1] for s in S])
    Tin = None if identity else np.stack(Tb)
    f64 = BlockTriFactor("rb").factorize(Tin, SiM, SiP)
    f32 = BlockTriFactor("rb").factorize_mixed(Tin, SiM, SiP)
    assert int((f32.info != 0).sum()) == 0
    B = T.tensor(np.random.default_rng(5).standard_normal((nb * m, r)), device="cuda")
    x64 = f64.solve(B)
    errs = []
    for iters in (0, 1, 2):
        X = T.empty_like(B)
        f32.solve_mixed_(B, X, iters=iters)
        errs.append((T.linalg.norm(X - x64) / T.linalg.norm(x64)).item())
    assert 1e-10 < errs[0] < 1e-4
    assert errs[1] < 1e-8
    assert errs[2] < 1e-10
    with pytest.raises(ValueError):
        f32.solve(B)


def test_c2_size_mixed_vs_oracle(T):
    """The C2 workload at full size (variable-coefficient Helmholtz kappa = 20, 32 slabs,
    p = 16, 32 leaves per interface: m = 512, 31 interfaces, 64 RHS of N(0,1) boundary
    data): the FP32 factorization refined with one FP64 residual step, and the FP64
    folded solve, against the oracle's FP64 Thomas solve (scipy lu_factor / lu_solve,
    oracle/blocktri.py) of the same S-system: mixed <= 1e-8, FP64 <= 1e-10."""
    from sslablu_amd.blocktri import BlockTriFactor
    from sslablu_amd.hps import HPSSlabOperator, SlabGrid2D
    from sslablu_amd.serving import ProblemSpec, make_pdo
    spec = ProblemSpec("helmholtz", 32, 16, 32)
    g = SlabGrid2D(spec.slabs, 1, spec.leaves, spec.p + 2)
    op = HPSSlabOperator(make_pdo(spec), g).build()
    nb, m = g.nb, g.m
    assert (nb, m) == (31, 512)
    rng = np.random.default_rng(0)
    F = T.tensor(rng.standard_normal((g.gb_points().shape[0], 64)), device="cuda")
    B = op.rhs_from_boundary(F)
    f32 = BlockTriFactor("rb").factorize_mixed(None, op.SiM, op.SiP)
    Xm = T.empty_like(B)
    f32.solve_mixed_(B.contiguous(), Xm, iters=1)
    f64 = BlockTriFactor("rb").factorize(None, op.SiM, op.SiP)
    X64 = op.fold(f64).solve(F)
    SiM, SiP = op.SiM.cpu().numpy(), op.SiP.cpu().numpy()
    S = [(np.zeros((m, m)) if i == 0 else SiM[i], np.zeros((m, m)) if i == nb - 1 else SiP[i])
         for i in range(nb)]
    x_ref = ob.thomas_direct(S, [np.eye(m)] * nb, B.cpu().numpy())
    nrm = np.linalg.norm(x_ref)
    assert np.linalg.norm(Xm.cpu().numpy() - x_ref) / nrm < 1e-8
    assert np.linalg.norm(X64.cpu().numpy() - x_ref) / nrm < 1e-10
    # the folded mixed solve (FP32 maps + FP32 backward sweep, one FP64 refinement step)
    f64.enable_mixed(None, op.SiM, op.SiP)
    fold = op.fold(f64).enable_mixed()
    Xf = T.empty_like(B)
    fold.solve_mixed_(F.contiguous(), Xf, iters=1)
    assert np.linalg.norm(Xf.cpu().numpy() - x_ref) / nrm < 1e-8


@pytest.mark.parametrize("pde,slabs,p", [("poisson", 8, 8), ("helmholtz", 16, 12),
                                         ("helmholtz", 5, 8)])
def test_folded_mixed_solve(T, pde, slabs, p):
    """sslab_oms_fold_solve_mixed against the FP64 folded solve (itself pinned to the
    oracle) of the same boundary data: without refinement the FP32 result is off by
    about the FP32 solve's error (between 1e-9 and 1e-3: FP32 arithmetic really ran),
    one step reaches <= 1e-8, two <= 1e-10; the fold's FP32 maps count in its bytes."""
    from sslablu_amd.blocktri import BlockTriFactor
    from sslablu_amd.hps import HPSSlabOperator, SlabGrid2D
    from sslablu_amd.serving import ProblemSpec, make_pdo
    spec = ProblemSpec(pde, slabs, p, 32)
    g = SlabGrid2D(spec.slabs, 1, spec.leaves, spec.p + 2)
    op = HPSSlabOperator(make_pdo(spec), g).build()
    fac = BlockTriFactor("rb").factorize(None, op.SiM, op.SiP)
    fold = op.fold(fac)
    b64 = fold.nbytes
    fac.enable_mixed(None, op.SiM, op.SiP)
    fold.enable_mixed()
    assert fold.nbytes == b64 + b64 // 2
    rng = np.random.default_rng(slabs + p)
    F = T.tensor(rng.standard_normal((g.gb_points().shape[0], 7)), device="cuda")
    X64 = fold.solve(F).cpu().numpy()
    nrm = np.linalg.norm(X64)
    work = fold.mixed_work(7)
    for iters, lo, hi in ((0, 1e-9, 1e-3), (1, 0.0, 1e-8), (2, 0.0, 1e-10)):
        X = T.full((g.nb * g.m, 7), float("nan"), dtype=T.float64, device="cuda")
        fold.solve_mixed_(F, X, iters=iters, work=work)
        err = np.linalg.norm(X.cpu().numpy() - X64) / nrm
        assert lo < err < hi, (iters, err)
