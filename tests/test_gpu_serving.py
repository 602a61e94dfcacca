"""GPU: continuous-batching serving.  Every request's result is the same alone and in
the mix (FP64: <= 1e-10 relative; mixed precision after refinement: <= 1e-8 against the
FP64 result), equals an eager FP64 solve of the same request (so graph replays, run many
times with eager work in between, match eager launches), and the factor cache admits,
evicts and rejects operators by its HBM budget."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = None


@pytest.fixture(scope="module")
def S(gpu):
    import torch
    from sslablu_amd import _lib, serving
    _lib.load()
    specs = [serving.ProblemSpec("poisson", 4, 6, 4), serving.ProblemSpec("helmholtz", 5, 8, 4),
             serving.ProblemSpec("poisson", 8, 4, 8), serving.ProblemSpec("helmholtz", 3, 10, 2)]
    return torch, serving, specs


def _rel(a, b):
    return (a - b).norm().item() / b.norm().item()


def _alone(serving, cache, reqs, **kw):
    srv = serving.BatchingServer(cache, **kw)
    out = []
    for q in reqs:
        q1 = serving.Request(q.rid, q.spec, q.F)
        srv.serve([q1], realtime=False)
        out.append(q1.X.clone())
    return out


def test_alone_equals_mix_fp64(S):
    torch, serving, specs = S
    cache = serving.FactorCache(max_operators=8)
    reqs = serving.make_requests(specs, 40, math.inf, rmax=6, seed=1)
    alone = _alone(serving, cache, reqs, slots=3, max_cols=16)
    srv = serving.BatchingServer(cache, slots=3, max_cols=16)
    st = srv.serve(reqs, realtime=False)
    assert st.n == 40 and st.batches < 40  # requests were batched together
    assert max(q.batch_size for q in reqs) > 1
    for q, xa in zip(reqs, alone):
        assert _rel(q.X, xa) < 1e-10
        e = cache.ops[q.spec]
        x_eager = e.fac.solve(e.op.rhs_from_boundary(q.F))
        assert _rel(q.X, x_eager) < 1e-10
    assert st.p99_ms >= st.p50_ms > 0


def test_realtime_arrivals_and_plain_path(S):
    torch, serving, specs = S
    cache = serving.FactorCache(max_operators=8)
    reqs = serving.make_requests(specs, 30, 3000.0, rmax=5, seed=2)
    srv = serving.BatchingServer(cache, slots=2, max_cols=32)
    srv.warm(specs)
    st = srv.serve(reqs)
    xs = [q.X.clone() for q in reqs]
    assert all(q.t_done >= q.t_arrive for q in reqs)
    plain = [serving.Request(q.rid, q.spec, q.F, q.t_arrive) for q in reqs]
    sp = serving.serve_plain(cache, plain)
    assert sp.batches == 30
    for x, q in zip(xs, plain):
        assert _rel(x, q.X) < 1e-10


def test_mixed_precision_serving(S):
    torch, serving, specs = S
    cache = serving.FactorCache(max_operators=8, mixed=True)
    reqs = serving.make_requests(specs, 24, math.inf, rmax=6, seed=3)
    alone = _alone(serving, cache, reqs, slots=2, max_cols=16)  # FP64, alone
    srv = serving.BatchingServer(cache, slots=2, max_cols=16, precision="mixed", iters=2)
    srv.serve(reqs, realtime=False)
    for q, xa in zip(reqs, alone):
        assert _rel(q.X, xa) < 1e-8
    # byte accounting of a mixed cache: the FP32 factor copy counted once, the fold's
    # FP32 maps in its bytes, both inside the admission estimate; FP64 and mixed batches
    # on this cache run the folded path
    for e in cache.ops.values():
        f = e.fac
        held = (f.factor.numel() * 8 + f.factor32.numel() * 4
                + (e.op.Y.numel() + e.op.SiM.numel() + e.op.SiP.numel()) * 8 + e.fold.nbytes)
        assert e.nbytes - e.graph_bytes == held
        assert held <= cache.estimate_bytes(e.spec)
        assert f._M64.data_ptr() == e.op.SiM.data_ptr()  # residual blocks are not copies
        prog = srv._program(e, 16, torch.cuda.current_stream())
        assert prog.V is None and prog.work is not None  # mixed: folded FP32 + refinement
        prog64 = serving._BatchProgram(e, 8, "fp64", 1, False, torch.cuda.current_stream())
        assert prog64.V is None      # FP64 batches on a mixed cache: the folded solve


def test_cache_eviction_and_budget(S):
    torch, serving, specs = S
    cache = serving.FactorCache(max_operators=2)
    reqs = serving.make_requests(specs[:3], 18, math.inf, rmax=4, seed=4)
    srv = serving.BatchingServer(cache, slots=1, max_cols=8)
    srv.serve(reqs, realtime=False)
    assert cache.evictions > 0 and len(cache.ops) <= 2
    for e in cache.ops.values():  # admission estimate bounds what a built operator holds
        assert e.fold is not None and e.fold.nbytes > 0
        assert e.nbytes - e.graph_bytes <= cache.estimate_bytes(e.spec)
    ref = serving.FactorCache(max_operators=8)
    for q in reqs:
        e = ref.acquire(q.spec)
        assert _rel(q.X, e.fac.solve(e.op.rhs_from_boundary(q.F))) < 1e-10
        ref.release(e)
    tiny = serving.FactorCache(hbm_budget_bytes=1024)
    assert tiny.acquire(specs[0]) is None
    with pytest.raises(RuntimeError):
        serving.BatchingServer(tiny).serve(
            serving.make_requests(specs[:1], 2, math.inf, rmax=2, seed=5), realtime=False)


def test_evicted_operator_is_freed_without_the_garbage_collector(S):
    """An evicted operator and its captured programs are released when the cache drops
    them, not later by the cyclic garbage collector (which may run inside another
    program's graph capture, where destroying a graph is not allowed): with the collector
    off, the evicted entry is gone right after the eviction."""
    import gc
    import weakref
    torch, serving, specs = S
    cache = serving.FactorCache(max_operators=1)
    srv = serving.BatchingServer(cache, slots=1, max_cols=8)
    e0 = cache.acquire(specs[0])
    srv._program(e0, 8, srv.streams[0])  # a captured program, kept in e0.graphs
    assert e0.graphs
    cache.release(e0)
    ref = weakref.ref(e0)
    del e0
    gc_was = gc.isenabled()
    gc.disable()
    try:
        e1 = cache.acquire(specs[1])  # evicts specs[0]
        assert cache.evictions == 1
        assert ref() is None
        cache.release(e1)
    finally:
        if gc_was:
            gc.enable()
