"""GPU: several block-tridiagonal systems of different sizes solved together (one
grouped GEMM launch per elimination level) equal their separate solves, and the
Gaussian-kernel test systems (T != I) match the FP64 oracle.  Tolerance 1e-10."""
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


def _random_system(nb, m, seed):
    rng = np.random.default_rng(seed)
    Tb = np.stack([np.eye(m) * 2 + 0.3 * rng.standard_normal((m, m)) / np.sqrt(m)
                   for _ in range(nb)])
    SiM = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiP = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    SiM[0] = 0
    SiP[-1] = 0
    return Tb, SiM, SiP


def test_solve_multi_matches_separate_and_oracle(T):
    from sslablu_amd.blocktri import BlockTriFactor, solve_multi
    cases = [(7, 64, 5), (12, 96, 3), (3, 128, 8), (1, 32, 2), (16, 48, 1), (5, 40, 0)]
    facs, Vs, Xs, refs, B0 = [], [], [], [], []
    for k, (nb, m, r) in enumerate(cases):
        Tb, SiM, SiP = _random_system(nb, m, k)
        f = BlockTriFactor("rb").factorize(Tb, SiM, SiP)
        B = np.random.default_rng(100 + k).standard_normal((nb * m, r))
        facs.append(f)
        Vs.append(T.tensor(B, device="cuda"))
        Xs.append(T.empty(nb * m, r, dtype=T.float64, device="cuda"))
        B0.append(B)
        S = [(SiM[i], SiP[i]) for i in range(nb)]
        refs.append(ob.thomas_direct(S, list(Tb), B) if r else np.zeros((nb * m, 0)))
    sep = [f.solve(T.tensor(B, device="cuda")) for f, B in zip(facs, B0)]
    solve_multi(facs, Vs, Xs)
    for X, Xs1, ref, (nb, m, r) in zip(Xs, sep, refs, cases):
        if r == 0:
            continue
        assert (T.linalg.norm(X - Xs1) / T.linalg.norm(Xs1)).item() < 1e-13
        assert np.linalg.norm(X.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-10


@pytest.mark.parametrize("nb,n", [(8, 128), (5, 96)])
def test_gaussian_blocktri_vs_oracle(T, nb, n):
    from sslablu_amd.blocktri import BlockTriFactor
    from sslablu_amd.synthetic import gaussian_blocktri, gaussian_points
    Td, Md, Pd = gaussian_blocktri(nb, n, seed=3, device="cuda")
    x = gaussian_points(nb, n, seed=3)
    K = lambda a, b: np.exp(-(a[:, None] - b[None, :]) ** 2 / 0.01)  # noqa: E731
    Tn = [K(x[b], x[b]) + n * np.eye(n) for b in range(nb)]
    S = [(0.1 * K(x[b], x[b - 1]) if b else np.zeros((n, n)),
          0.1 * K(x[b], x[b + 1]) if b < nb - 1 else np.zeros((n, n))) for b in range(nb)]
    assert np.abs(Td.cpu().numpy() - np.stack(Tn)).max() < 1e-12
    rng = np.random.default_rng(1)
    B = rng.standard_normal((nb * n, 4))
    ref = ob.thomas_direct(S, Tn, B)
    for method in ("rb", "thomas"):
        f = BlockTriFactor(method).factorize(Td, Md, Pd)
        x_ = f.solve(T.tensor(B, device="cuda")).cpu().numpy()
        assert np.linalg.norm(x_ - ref) / np.linalg.norm(ref) < 1e-10


def test_solve_multi_routes_wide_systems(T):
    """Systems with more than 32 right-hand sides take their own launches while the
    others share one (here under a small workgroup target, so the grouped launches split
    K): every system still matches its separate solve and the oracle."""
    from sslablu_amd import tuning
    from sslablu_amd.blocktri import BlockTriFactor, solve_multi
    cases = [(17, 256, 96), (9, 128, 80), (7, 64, 5), (12, 96, 33), (4, 32, 2)]
    pol = tuning.Policy(16, 64, 8, 64, 3, 3, 0)
    facs, Vs, Xs, refs, B0 = [], [], [], [], []
    for k, (nb, m, r) in enumerate(cases):
        Tb, SiM, SiP = _random_system(nb, m, 50 + k)
        f = BlockTriFactor("rb").factorize(Tb, SiM, SiP)
        B = np.random.default_rng(200 + k).standard_normal((nb * m, r))
        facs.append(f)
        Vs.append(T.tensor(B, device="cuda"))
        Xs.append(T.empty(nb * m, r, dtype=T.float64, device="cuda"))
        B0.append(B)
        S = [(SiM[i], SiP[i]) for i in range(nb)]
        refs.append(ob.thomas_direct(S, list(Tb), B))
    with tuning.policy(pol):
        sep = [f.solve(T.tensor(B, device="cuda")) for f, B in zip(facs, B0)]
        solve_multi(facs, Vs, Xs)
    for X, Xs1, ref in zip(Xs, sep, refs):
        assert (T.linalg.norm(X - Xs1) / T.linalg.norm(Xs1)).item() < 1e-13
        assert np.linalg.norm(X.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-10
