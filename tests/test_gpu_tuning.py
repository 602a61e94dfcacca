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


@pytest.mark.parametrize("cfg", [10, 11, 12, 13, 14])
def test_register_streamed_tiles(T, cfg):
    """Tiles 10 / 11 stream A from HBM into MFMA registers (aligned launches with at most
    64 columns), tiles 12-14 take 32- or 64-deep K slabs: GEMMs with ragged M, K tails and
    split-K, and a red-black solve, match."""
    from sslablu_amd import tuning
    from sslablu_amd._lib import call, ptr, stream_handle
    from sslablu_amd.blocktri import BlockTriFactor
    pol = tuning.Policy(100000, 1024, 64, 64, cfg, cfg, 0)
    rng = np.random.default_rng(cfg)
    for (b, M, N, K) in [(3, 200, 64, 1000), (2, 512, 48, 96), (1, 70, 64, 34), (5, 128, 2, 512)]:
        A = rng.standard_normal((b, M, K))
        Bm = rng.standard_normal((b, K, N))
        C0 = rng.standard_normal((b, M, N))
        At, Bt, Ct = (T.tensor(x, device="cuda") for x in (A, Bm, C0))
        with tuning.policy(pol):
            call("sslab_dgemm_strided_batched", 0, 0, M, N, K, -0.5, ptr(At), K, M * K, ptr(Bt), N,
                 K * N, 1.0, ptr(Ct), N, M * N, b, stream_handle())
        ref = -0.5 * (A @ Bm) + C0
        assert np.linalg.norm(Ct.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-13
    nb, m = 13, 96
    SiM = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiP = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    fac = BlockTriFactor("rb").factorize(None, SiM, SiP)
    Bv = T.tensor(rng.standard_normal((nb * m, 64)), device="cuda")
    ref = fac.solve(Bv)
    with tuning.policy(pol):
        x = fac.solve(Bv)
    assert (T.linalg.norm(x - ref) / T.linalg.norm(ref)).item() < 1e-12


def test_backward_level_policies(T):
    """Per-level tile policies of the red-black backward sweep (sslab_rb_set_backward_policies)
    change speed only: the same solution under a mixed table, invalid entries are rejected,
    and the per-level autotuner's table is used by later solves of that width."""
    from sslablu_amd import tuning
    from sslablu_amd._lib import SslabError
    from sslablu_amd.blocktri import BlockTriFactor, rb_levels
    rng = np.random.default_rng(5)
    nb, m, r = 31, 64, 32
    SiM = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiP = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    fac = BlockTriFactor("rb").factorize(None, SiM, SiP)
    B = T.tensor(rng.standard_normal((nb * m, r)), device="cuda")
    ref = fac.solve(B)
    levels, _ = rb_levels(nb, m)
    assert levels >= 2
    table = [tuning.Policy(100000, 4096, 64, 64, 5, 5, 0), tuning.Policy(1, 512, 64, 1 << 30, 4, 4, 2),
             tuning.DEFAULT] * levels
    with tuning.backward_policies(table[:levels]):
        x = fac.solve(B)
    assert (T.linalg.norm(x - ref) / T.linalg.norm(ref)).item() < 1e-12
    with pytest.raises(SslabError):
        tuning.set_backward_policies([tuning.Policy(cfg_split=99)])
    best, us = fac.autotune_backward_steps(r, reps=2, cands=tuning.candidates()[::40])
    assert len(best) == levels and us > 0 and fac._bwd_policies[r] == best
    x = fac.solve(B)
    assert (T.linalg.norm(x - ref) / T.linalg.norm(ref)).item() < 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("cfg", [3, 4, 5, 12])
def test_split_k_outputs_initialised_together(T, cfg):
    """A grouped launch forced to split K: its non-in-place outputs start as beta Cin (or
    zero without Cin) through one initialisation launch for all descriptors (different
    shapes, batches, leading dimensions, betas), then the K parts add atomically; an
    in-place descriptor (Cin == C) accumulates onto its own values.  Against numpy."""
    from sslablu_amd import tuning
    from sslablu_amd.hbs import gdesc, ggemm
    rng = np.random.default_rng(cfg)
    dev = "cuda"
    cases = []
    for nb, M, N, K, beta, kind in ((3, 70, 40, 900, 0.5, "cin"), (1, 33, 64, 1200, 0.0, "none"),
                                    (5, 128, 24, 700, -2.0, "cin"), (2, 50, 16, 1500, 1.0, "inplace")):
        A = rng.standard_normal((nb, M, K))
        B = rng.standard_normal((nb, K, N))
        C0 = rng.standard_normal((nb, M, N))
        ref = A @ B + (beta * C0 if kind != "none" else 0.0)
        tA, tB = T.tensor(A, device=dev), T.tensor(B, device=dev)
        if kind == "inplace":
            out = T.tensor(C0, device=dev)
            d = gdesc(out, [(tA, tB, 1.0)], cin=out, beta=beta)
        else:
            out = T.full((nb, M, N), float("nan"), dtype=T.float64, device=dev)
            cin = T.tensor(C0, device=dev) if kind == "cin" else None
            d = gdesc(out, [(tA, tB, 1.0)], cin=cin, beta=beta)
        cases.append((d, out, ref))
    # a target no tile reaches and a small split threshold: every descriptor splits K
    with tuning.policy(tuning.Policy(1 << 30, 4096, 16, 64, cfg, cfg, 0)):
        ggemm(*[d for d, _, _ in cases])
    T.cuda.synchronize()
    for _, out, ref in cases:
        assert rel(out.cpu().numpy(), ref) < 1e-13
