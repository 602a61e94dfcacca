"""GPU: the GEMM tile policy and the autotuner change speed only, never results
beyond summation order (FP64 solves agree to 1e-12 under every policy)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


def test_policy_roundtrip_and_validation(T):
    from sslablu_amd import tuning
    from sslablu_amd._lib import SslabError
    old = tuning.get_policy()
    p = tuning.Policy(512, 2048, 16, 128, 2, 6, 1)
    with tuning.policy(p):
        assert tuning.get_policy() == p
    assert tuning.get_policy() == old
    with pytest.raises(SslabError):
        tuning.set_policy(tuning.Policy(cfg_mid=17))


@pytest.mark.parametrize("r", [1, 7, 64])
def test_solve_same_under_all_policies(T, r):
    from sslablu_amd import tuning
    from sslablu_amd.blocktri import BlockTriFactor
    rng = np.random.default_rng(r)
    nb, m = 9, 96
    SiM = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiP = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    fac = BlockTriFactor("rb").factorize(np.stack([np.eye(m)] * nb), SiM, SiP)
    B = T.tensor(rng.standard_normal((nb * m, r)), device="cuda")
    with tuning.policy(tuning.DEFAULT):
        ref = fac.solve(B)
    for p in tuning.candidates()[::7]:
        with tuning.policy(p):
            x = fac.solve(B)
        assert (T.linalg.norm(x - ref) / T.linalg.norm(ref)).item() < 1e-12
    pol, us = fac.autotune(r, reps=2)
    assert tuning.lookup(fac.tune_key(r)) == pol and us > 0
    x = fac.solve(B)
    assert (T.linalg.norm(x - ref) / T.linalg.norm(ref)).item() < 1e-12
