"""GPU: the allocation-free mixed-precision solve (sslab_rb_solve_mixed_noalloc) against
the workspace entry point (sslab_rb_solve_mixed) through the C ABI: two calls in a row,
the second with a smaller input (the scratch keeps the first call's contents), each equal
to the workspace result (same kernels on the same data: 1e-14 relative) and within 1e-8
of the FP64 solve; an input too large for the scratch is refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(torch, nb, m, seed):
    from sslablu_amd.blocktri import BlockTriFactor
    rng = np.random.default_rng(seed)
    T = np.eye(m)[None] + 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    M = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    P = 0.3 * rng.standard_normal((nb, m, m)) / np.sqrt(m)
    dev = lambda a: torch.tensor(a, device="cuda")  # noqa: E731
    f64 = BlockTriFactor("rb").factorize(dev(T), dev(M), dev(P))
    f32 = BlockTriFactor("rb").factorize_mixed(dev(T), dev(M), dev(P))
    return f64, f32


def _call(torch, name, f32, B, X, iters, work=None):
    from sslablu_amd._lib import call, ptr
    args = [f32.nb, f32.m, B.shape[1], ptr(f32.factor32), ptr(f32._T64), ptr(f32._M64),
            ptr(f32._P64), ptr(B), ptr(X)]
    if work is not None:
        args += [ptr(work[0]), ptr(work[1])]
    call(name, *args, iters, None)


def test_noalloc_matches_workspace_entry_point(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    rng = np.random.default_rng(0)
    for nb, m, r, seed in ((15, 160, 24, 1), (7, 96, 5, 2)):  # second call: smaller input
        f64, f32 = _system(torch, nb, m, seed)
        N = nb * m
        B = torch.tensor(rng.standard_normal((N, r)), device="cuda")
        Xw = torch.empty_like(B)
        Xn = torch.full_like(B, float("nan"))
        work = (torch.empty(N * r, dtype=torch.float64, device="cuda"),
                torch.empty(2 * N * r, dtype=torch.float32, device="cuda"))
        _call(torch, "sslab_rb_solve_mixed", f32, B, Xw, 2, work)
        _call(torch, "sslab_rb_solve_mixed_noalloc", f32, B, Xn, 2)
        x64 = f64.solve(B)
        torch.cuda.synchronize()
        assert ((Xn - Xw).norm() / Xw.norm()).item() < 1e-14
        assert ((Xn - x64).norm() / x64.norm()).item() < 1e-8
        # the Python wrapper takes the same route
        Xp = f32.solve_mixed_noalloc_(B, torch.empty_like(B), iters=2)
        assert ((Xp - Xw).norm() / Xw.norm()).item() < 1e-14


def test_noalloc_refuses_oversized_input(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    f64, f32 = _system(torch, 4, 64, 3)
    r = (64 << 20) // (16 * 4 * 64) + 1           # one column more than the scratch holds
    B = torch.zeros(4 * 64, r, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.SslabError, match="scratch"):
        _call(torch, "sslab_rb_solve_mixed_noalloc", f32, B, torch.empty_like(B), 1)
