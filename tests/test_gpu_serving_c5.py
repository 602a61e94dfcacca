"""GPU: the serving workload at its stated scale -- 24 operators ({Poisson, Helmholtz} x
slabs {8, 16, 32, 64} x leaf order p in {8, 12, 16}, 32 leaves per interface) and 2000
requests with r ~ U[1, 32] columns of N(0,1) boundary data, all admitted by one mixed
cache (FP32 factor copies + folded FP64 maps) under an HBM budget.

Checks, per request: the batched FP64 result equals the same request solved alone
(<= 1e-10 relative), the batched mixed-precision result (FP32 solve + one FP64
refinement step) is within 1e-8 of it; the mixed cache's byte accounting (every tensor
counted once, inside the admission estimate); and a sample of requests on the smallest
operators of both PDEs against the oracle's FP64 oms solve (oracle/hps.py, 1e-10)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return (a - b).norm().item() / b.norm().item()


def test_c5_scale_alone_vs_mix_and_oracle(gpu):
    import torch
    from sslablu_amd import _lib, serving as sv
    _lib.load()
    specs = [sv.ProblemSpec(pde, n, p) for pde in ("poisson", "helmholtz")
             for n in (8, 16, 32, 64) for p in (8, 12, 16)]
    cache = sv.FactorCache(hbm_budget_bytes=64 << 30, max_operators=24, mixed=True)
    reqs = sv.make_requests(specs, 2000, math.inf, rmax=32, seed=11)
    print(f"[c5] {len(reqs)} requests over {len(specs)} operators", flush=True)
    srv64 = sv.BatchingServer(cache, slots=4, max_cols=128, precision="fp64")
    st = srv64.serve(reqs, realtime=False)
    assert st.n == 2000 and st.batches < 2000 and len(cache.ops) == 24
    X64 = [q.X.clone() for q in reqs]
    # byte accounting of the mixed cache at this scale
    for e in cache.ops.values():
        f = e.fac
        held = (f.factor.numel() * 8 + f.factor32.numel() * 4 + e.fold.nbytes
                + (e.op.Y.numel() + e.op.SiM.numel() + e.op.SiP.numel()) * 8)
        assert e.nbytes - e.graph_bytes == held <= cache.estimate_bytes(e.spec)
    assert cache.nbytes <= cache.budget
    # every request alone: the operator's folded FP64 solve of that request only
    for q, x in zip(reqs, X64):
        xa = cache.ops[q.spec].fold.solve(q.F)
        assert _rel(x, xa) < 1e-10, q.rid
    print("[c5] FP64 batched == alone for all requests", flush=True)
    srvmx = sv.BatchingServer(cache, slots=4, max_cols=128, precision="mixed", iters=1)
    mreqs = [sv.Request(q.rid, q.spec, q.F) for q in reqs]
    srvmx.serve(mreqs, realtime=False)
    errs = [_rel(q.X, x) for q, x in zip(mreqs, X64)]
    print(f"[c5] mixed max rel err vs FP64 {max(errs):.2e}", flush=True)
    assert max(errs) < 1e-8
    # a sample against the oracle (smallest operators, both PDEs)
    from oracle import hps as oh
    for pde in ("poisson", "helmholtz"):
        spec = sv.ProblemSpec(pde, 8, 8)
        q = next(r for r in reqs if r.spec == spec)
        e = cache.ops[spec]
        g = e.op.g
        P = g.gb_points()
        F = q.F.cpu().numpy()
        key = {tuple(np.round(pt, 12)): k for k, pt in enumerate(P)}

        def bc(Pts, c):
            return np.array([F[key[tuple(np.round(pt, 12))], c] for pt in Pts])

        pdo = sv.make_pdo(spec)
        for c in range(min(2, q.r)):
            ref = oh.oms_stot(pdo, spec.slabs, (g.hx, g.hy), g.q, bc=lambda Pts: bc(Pts, c))
            u = np.linalg.solve(oh.stot_dense(ref["S_rk_list"]), np.concatenate(ref["rhs_list"]))
            got = X64[q.rid][:, c].cpu().numpy()
            assert np.linalg.norm(got - u) / np.linalg.norm(u) < 1e-10, (pde, c)
