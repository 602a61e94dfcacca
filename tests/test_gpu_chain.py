"""GPU: the red-black backward sweep as one dependency-ordered launch
(SSLAB_RB_CHAINED_BACKWARD: every level's tiles in one grid, each tile starting once the
blocks it reads are final) against the level-by-level sweep and the oracle.

* Whole solves (general and identity diagonal blocks, 2 to 31 blocks, 1 to 64 columns):
  the chained solve equals the per-level one to 1e-13 and the oracle's Thomas solve to
  1e-10; more than 128 blocks fall back to per-level launches.
* Split phases and twenty chained solves in a row (the kernel resets its counters after
  every launch), each equal to the first.
* The folded boundary solve of a 2-D HPS operator with the chained sweep.
After each group the abort flag is read: no bounded wait may have run out."""
import numpy as np
import pytest

from oracle import blocktri as ob
from tests.test_oracle_hps import helmholtz

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


def _system(nb, m, ident, seed):
    rng = np.random.default_rng(seed)
    Tb = None if ident else np.stack([np.eye(m) + 0.3 * rng.standard_normal((m, m)) / np.sqrt(m)
                                      for _ in range(nb)])
    SiM = 0.4 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiP = 0.4 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiM[0] = 0
    SiP[-1] = 0
    return Tb, SiM, SiP, rng


@pytest.mark.parametrize("nb,m,r,ident", [(2, 16, 3, False), (5, 16, 3, True), (7, 160, 1, False),
                                          (12, 33, 5, True), (31, 128, 8, False),
                                          (31, 512, 64, True), (100, 24, 4, False),
                                          (130, 8, 2, True)])
def test_chained_equals_per_level_and_oracle(T, nb, m, r, ident):
    from sslablu_amd.blocktri import BlockTriFactor, chain_status
    print(f"[chain] nb={nb} m={m} r={r} ident={ident}", flush=True)
    Tb, SiM, SiP, rng = _system(nb, m, ident, nb * 7 + m)
    B = T.tensor(rng.standard_normal((nb * m, r)), device="cuda")
    fac = BlockTriFactor("rb").factorize(Tb, SiM, SiP)
    X0 = fac.solve(B)
    fac.chained_backward = True
    X1 = fac.solve(B)
    assert not chain_status()
    assert rel(X1.cpu().numpy(), X0.cpu().numpy()) < 1e-13
    Tl = [np.eye(m)] * nb if ident else list(Tb)
    S = [(SiM[i] if i > 0 else np.zeros((m, m)), SiP[i] if i < nb - 1 else np.zeros((m, m)))
         for i in range(nb)]
    x_ref = ob.thomas_direct(S, Tl, B.cpu().numpy())
    assert rel(X1.cpu().numpy(), x_ref) < 1e-10


def test_split_phases_and_repeated_launches(T):
    from sslablu_amd.blocktri import BlockTriFactor, chain_status
    nb, m, r = 31, 96, 16
    Tb, SiM, SiP, rng = _system(nb, m, False, 5)
    B = T.tensor(rng.standard_normal((nb * m, r)), device="cuda")
    fac = BlockTriFactor("rb").factorize(Tb, SiM, SiP)
    X0 = fac.solve(B)
    fac.chained_backward = True
    X1 = T.empty_like(B)
    fac.solve_(B.clone(), X1, phase="forward")
    fac.solve_(None, X1, phase="backward")
    assert rel(X1.cpu().numpy(), X0.cpu().numpy()) < 1e-13
    outs = [fac.solve(B) for _ in range(20)]
    assert not chain_status()
    for X in outs:
        assert rel(X.cpu().numpy(), X0.cpu().numpy()) < 1e-13


@pytest.mark.parametrize("N,nx,ny,q,r", [(9, 2, 2, 6, 4), (32, 1, 4, 8, 16)])
def test_folded_solve_with_chained_sweep(T, N, nx, ny, q, r):
    from sslablu_amd.blocktri import BlockTriFactor, chain_status
    from sslablu_amd.hps import HPSSlabOperator, SlabGrid2D
    op = HPSSlabOperator(helmholtz(6.0), SlabGrid2D(N, nx, ny, q)).build()
    F = T.tensor(np.random.default_rng(N).standard_normal((op.g.gb_points().shape[0], r)),
                 device=op.device)
    fac = BlockTriFactor("rb", op.device).factorize(None, op.SiM, op.SiP)
    fold = op.fold(fac)
    X0 = fold.solve(F)
    fac.chained_backward = True
    X1 = fold.solve(F)
    assert not chain_status()
    assert (T.linalg.norm(X1 - X0) / T.linalg.norm(X0)).item() < 1e-13
