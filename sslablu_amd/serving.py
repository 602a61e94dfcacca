"""Serving: a request loop over factorized slab operators resident on one GPU.

* ``FactorCache`` keeps built operators (HPS S-system maps + red-black factor, and the
  FP32 copy for mixed-precision solves) in HBM, admitted by a byte budget and an
  operator-count cap, evicting the least recently used operator that no batch in
  flight is using.
* ``BatchingServer`` does continuous batching: requests that arrive while others are
  on the GPU queue per operator; whenever one of its stream slots is free it launches
  the oldest operator's queued requests as one multi-column batch (boundary data ->
  rhs maps -> red-black solve), so different operators' batches overlap on the device
  and one operator's requests share its triangular sweeps.  Each (operator, batch
  width) pair is captured once as a HIP graph and replayed.
* ``serve_plain`` is the plain FP64 path for comparison: one request at a time, in
  arrival order, eager launches.

A request's result does not depend on what it was batched with: columns of a batch are
independent, so a result differs from the same request solved alone only by
floating-point summation order inside the GEMMs (tile / split-K choice for the batch
width), far below the stated tolerances (1e-10 FP64, 1e-8 mixed).

Request data (boundary values) and results stay on the device; latencies are host wall
times from a request's arrival to the completion of the batch that carried it.
"""
from __future__ import annotations

import gc
import math
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .blocktri import BlockTriFactor, rb_levels
from .hps import HPSSlabOperator, SlabGrid2D

F64 = torch.float64


# ---------------------------------------------------------------------------
# problems

@dataclass(frozen=True)
class ProblemSpec:
    """One slab operator: ``pde`` in {'poisson', 'helmholtz'} on the unit square with
    ``slabs`` single slabs (slabs - 1 interfaces), leaf order ``p`` and ``leaves`` leaves
    per interface (m = leaves * p interface nodes)."""
    pde: str
    slabs: int
    p: int
    leaves: int = 32
    kappa: float = 20.0

    @property
    def m(self):
        return self.leaves * self.p

    @property
    def nb(self):
        return self.slabs - 1


class _PDO:
    def __init__(self, c11, c22, c=None):
        self.c11, self.c22, self.c = c11, c22, c
        self.c1 = self.c2 = self.c12 = None


def make_pdo(spec: ProblemSpec):
    """-Lap u = 0, or -Lap u - kappa^2 (1 - 0.5 exp(-|x - (1/2,1/2)|^2 / 0.05)) u = 0."""
    one = lambda P: np.ones(P.shape[0])  # noqa: E731
    if spec.pde == "poisson":
        return _PDO(one, one)
    if spec.pde == "helmholtz":
        k2 = spec.kappa ** 2

        def c(P):
            r2 = (P[:, 0] - 0.5) ** 2 + (P[:, 1] - 0.5) ** 2
            return -k2 * (1.0 - 0.5 * np.exp(-r2 / 0.05))

        return _PDO(one, one, c)
    raise ValueError(f"unknown pde {spec.pde!r}")


class CachedOperator:
    """A built operator: rhs maps (op.Y), red-black factor, the boundary-folded forward
    sweep (FP64 batches: one launch + the backward sweep; built in every cache, so FP64
    batches on a mixed cache are folded too), and in mixed caches the FP32 factor and
    FP32 fold maps (mixed batches: the folded solve in FP32, then FP64 refinement)."""

    def __init__(self, spec: ProblemSpec, device, mixed: bool):
        t0 = time.perf_counter()
        g = SlabGrid2D(spec.slabs, 1, spec.leaves, spec.p + 2)
        op = HPSSlabOperator(make_pdo(spec), g, device).build()
        op.Ts = None  # slab DtN maps are not needed after assembly
        nb, m = g.nb, g.m
        fac = BlockTriFactor("rb", device).factorize(None, op.SiM, op.SiP)  # T = I
        if mixed:
            fac.enable_mixed(None, op.SiM, op.SiP)
        self.fold = op.fold(fac)
        if mixed:
            self.fold.enable_mixed()
        torch.cuda.synchronize(device)
        self.spec, self.op, self.fac = spec, op, fac
        self.N = nb * m
        self.ngb = g.gb_points().shape[0]
        self.factor_s = time.perf_counter() - t0
        self.pins = 0
        self.busy = False  # a batch of this operator is in flight (its buffers are in use)
        self.graphs: dict = {}
        self.graph_bytes = 0

    @property
    def nbytes(self):
        # fac.nbytes covers the FP64 factor and, in mixed caches, its FP32 copy; the FP64
        # blocks the refinement residual reads are op.SiM / op.SiP themselves (no copy);
        # fold.nbytes counts its FP32 maps in mixed caches
        n = self.fac.nbytes + (self.op.Y.numel() + self.op.SiM.numel() + self.op.SiP.numel()) * 8
        if self.fold is not None:
            n += self.fold.nbytes
        return n + self.graph_bytes


class FactorCache:
    """Operators resident in HBM, admitted by byte budget and count cap (LRU eviction of
    operators with no batch in flight)."""

    def __init__(self, device=None, hbm_budget_bytes: int = 64 << 30, max_operators: int = 24,
                 mixed: bool = False):
        _lib.require_gpu()
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.budget = int(hbm_budget_bytes)
        self.max_ops = int(max_operators)
        self.mixed = mixed
        self.ops: OrderedDict[ProblemSpec, CachedOperator] = OrderedDict()
        self.hits = self.misses = self.evictions = 0
        self.factor_s = 0.0

    @property
    def nbytes(self):
        return sum(e.nbytes for e in self.ops.values())

    def estimate_bytes(self, spec: ProblemSpec) -> int:
        """Resident bytes of a built operator before graphs: FP64 factor, its FP32 copy in
        mixed caches, the maps Y and coupling blocks, and an upper bound of the folded
        boundary maps (an odd node of level l reads at most 2^(l+2) strip windows of
        bottom and top data plus both boundary segments, a coarsest-level node all nF
        boundary values; 12 bytes per element in mixed caches, which keep them in FP32
        too)."""
        nb, m, p = spec.nb, spec.m, spec.p
        ncols = 2 * m + 4 * p
        rb, _, _ = _lib.sizes("rb", nb, m)
        n = 8 * (rb + nb * m * ncols + 2 * nb * m * m)
        if self.mixed:
            n += 4 * rb
        L, n_top = rb_levels(nb, m)
        nF = 2 * spec.slabs * p + 2 * m
        cols, k = n_top * nF, nb
        for lvl in range(L):
            cols += (k // 2) * min(nF, (4 << lvl) * 2 * p + 2 * m)
            k = (k + 1) // 2
        n += (12 if self.mixed else 8) * m * cols
        return n

    def _evict_for(self, need: int) -> bool:
        while (self.nbytes + need > self.budget or len(self.ops) >= self.max_ops) and self.ops:
            victim = next((s for s, e in self.ops.items() if e.pins == 0), None)
            if victim is None:
                return False
            e = self.ops.pop(victim)
            # its captured programs go now, outside any capture (a graph destroyed while
            # another stream is capturing aborts the process)
            e.graphs.clear()
            self.evictions += 1
        return self.nbytes + need <= self.budget and len(self.ops) < self.max_ops

    def acquire(self, spec: ProblemSpec) -> CachedOperator | None:
        """The operator for `spec`, building it if needed; None when it cannot be admitted
        now (budget taken by operators that batches in flight still use)."""
        e = self.ops.get(spec)
        if e is not None:
            self.hits += 1
            self.ops.move_to_end(spec)
            e.pins += 1
            return e
        if not self._evict_for(self.estimate_bytes(spec)):
            return None
        self.misses += 1
        e = CachedOperator(spec, self.device, self.mixed)
        self.factor_s += e.factor_s
        self.ops[spec] = e
        e.pins += 1
        return e

    def release(self, e: CachedOperator):
        e.pins -= 1


# ---------------------------------------------------------------------------
# requests

@dataclass
class Request:
    rid: int
    spec: ProblemSpec
    F: torch.Tensor               # (ngb, r) float64 boundary data on the device
    t_arrive: float = 0.0         # seconds after the start of serving
    X: torch.Tensor | None = None  # (N, r) interface solution on the device
    t_done: float = math.nan
    batch_size: int = 0

    @property
    def r(self):
        return self.F.shape[1]

    @property
    def latency(self):
        return self.t_done - self.t_arrive


def make_requests(specs, n, rate, rmax=32, seed=0, device=None, ngb_of=None):
    """`n` requests with Poisson arrivals at `rate` per second (rate = inf: all at 0),
    each uniform over `specs` with r uniform in [1, rmax] and N(0,1) boundary data."""
    rng = np.random.default_rng(seed)
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    t = 0.0
    out = []
    for i in range(n):
        spec = specs[rng.integers(len(specs))]
        r = int(rng.integers(1, rmax + 1))
        ngb = (ngb_of or gb_rows)(spec)
        F = torch.tensor(rng.standard_normal((ngb, r)), device=dev)
        if math.isfinite(rate):
            t += rng.exponential(1.0 / rate)
        out.append(Request(i, spec, F, t))
    return out


def gb_rows(spec: ProblemSpec) -> int:
    """Rows of a request's boundary data for `spec`."""
    return 2 * spec.slabs * spec.p + 2 * spec.m


# ---------------------------------------------------------------------------
# the batched solve of one operator, optionally captured as a HIP graph

class _BatchProgram:
    """Static buffers + (optionally) a captured graph: Fb (ngb x W) -> V -> Xb (N x W)."""

    def __init__(self, e: CachedOperator, W: int, precision: str, iters: int, use_graph: bool,
                 stream):
        dev = e.op.device
        # the operator's parts, not the CachedOperator itself: the operator keeps its
        # programs (e.graphs), and a reference back would make a cycle that only the
        # garbage collector frees -- possibly in the middle of another program's capture
        self.op, self.fac, self.fold = e.op, e.fac, e.fold
        self.W, self.precision, self.iters = W, precision, iters
        self.Fb = torch.zeros(e.ngb, W, dtype=F64, device=dev)
        folded = e.fold is not None
        self.V = None if folded else torch.empty(e.N, W, dtype=F64, device=dev)
        self.Xb = torch.empty(e.N, W, dtype=F64, device=dev)
        self.work = None
        if precision == "mixed":
            # folded: FP64 residual, FP32 panels (+ boundary data) and the FP64 rhs
            self.work = (e.fold.mixed_work(W) if folded else
                         (torch.empty(e.N * W, dtype=F64, device=dev),
                          torch.empty(2 * e.N * W, dtype=torch.float32, device=dev)))
        self.nbytes = (self.Fb.numel() + (1 if folded else 2) * e.N * W) * 8 + (
            sum(t.numel() * t.element_size() for t in self.work) if self.work is not None else 0)
        self.graph = None
        with torch.cuda.stream(stream):
            self._run()  # eager once: code objects, tuned policy
        stream.synchronize()
        if use_graph:
            g = torch.cuda.CUDAGraph()
            # no garbage collection while capturing: freeing unrelated cyclic garbage (a
            # graph, pinned memory) on another stream during the capture is not allowed
            gc_was = gc.isenabled()
            gc.disable()
            try:
                with torch.cuda.graph(g, stream=stream):
                    self._run()
            finally:
                if gc_was:
                    gc.enable()
            stream.synchronize()
            with torch.cuda.stream(stream):
                g.replay()  # the first replay uploads the graph: pay it here, not in a batch
            stream.synchronize()
            self.graph = g

    def _run(self):
        if self.V is None:  # folded forward sweep, then the backward sweep
            if self.precision == "mixed":
                self.fold.solve_mixed_(self.Fb, self.Xb, iters=self.iters, work=self.work)
            else:
                self.fold.solve_(self.Fb, self.Xb)
            return
        self.op.rhs_from_boundary(self.Fb, out=self.V)
        if self.precision == "mixed":
            self.fac.solve_mixed_(self.V, self.Xb, iters=self.iters, work=self.work)
        else:
            self.fac.solve_(self.V, self.Xb)

    def launch(self):
        if self.graph is not None:
            self.graph.replay()
        else:
            self._run()


def _bucket(R, buckets):
    for b in buckets:
        if R <= b:
            return b
    return R


@dataclass
class _Batch:
    slot: int
    entry: CachedOperator
    reqs: list
    event: torch.cuda.Event
    t_launch: float


@dataclass
class ServeStats:
    n: int = 0
    rhs: int = 0
    wall_s: float = 0.0
    batches: int = 0
    mean_batch_requests: float = 0.0
    p50_ms: float = 0.0
    p99_ms: float = 0.0
    mean_ms: float = 0.0
    solves_per_s: float = 0.0
    requests_per_s: float = 0.0
    extra: dict = field(default_factory=dict)


def _stats(reqs, wall, batches):
    lat = np.array([q.latency for q in reqs]) * 1e3
    rhs = sum(q.r for q in reqs)
    return ServeStats(n=len(reqs), rhs=rhs, wall_s=wall, batches=batches,
                      mean_batch_requests=len(reqs) / max(batches, 1),
                      p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                      mean_ms=float(lat.mean()), solves_per_s=rhs / wall,
                      requests_per_s=len(reqs) / wall)


class BatchingServer:
    """Continuous-batching request loop over a FactorCache (one host thread)."""

    def __init__(self, cache: FactorCache, slots: int = 4, max_cols: int = 128,
                 precision: str = "fp64", iters: int = 1, use_graphs: bool = True,
                 buckets=(8, 16, 32, 64, 128, 256)):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision is 'fp64' or 'mixed'")
        if precision == "mixed" and not cache.mixed:
            raise ValueError("mixed-precision serving needs FactorCache(mixed=True)")
        self.cache = cache
        self.precision = precision
        self.iters = iters
        self.use_graphs = use_graphs
        self.max_cols = max_cols
        self.buckets = tuple(b for b in buckets if b <= max_cols) + (max_cols,)
        self.streams = [torch.cuda.Stream(device=cache.device) for _ in range(slots)]

    def _program(self, e: CachedOperator, W: int, stream) -> _BatchProgram:
        key = (W, self.precision, self.iters)
        prog = e.graphs.get(key)
        if prog is None:
            prog = _BatchProgram(e, W, self.precision, self.iters, self.use_graphs, stream)
            e.graphs[key] = prog
            e.graph_bytes += prog.nbytes
        return prog

    def warm(self, specs, widths=None):
        """Build operators and capture their batch programs ahead of serving."""
        widths = widths or self.buckets
        s = self.streams[0]
        for spec in specs:
            e = self.cache.acquire(spec)
            if e is None:
                raise RuntimeError("HBM budget too small to hold the warm set")
            for W in widths:
                self._program(e, W, s)
            self.cache.release(e)
        torch.cuda.synchronize(self.cache.device)

    def _launch(self, slot, e, reqs, t_now):
        s = self.streams[slot]
        e.busy = True
        R = sum(q.r for q in reqs)
        W = _bucket(R, self.buckets)
        prog = self._program(e, W, s)
        with torch.cuda.stream(s):
            # one gather in and one copy out per batch (columns past R keep stale data:
            # every column is solved independently and those results are not read)
            if len(reqs) == 1:
                prog.Fb[:, :R].copy_(reqs[0].F)
            else:
                prog.Fb[:, :R].copy_(torch.cat([q.F for q in reqs], dim=1))
            prog.launch()
            Xall = prog.Xb[:, :R].clone()
            c = 0
            for q in reqs:
                q.X = Xall[:, c:c + q.r]
                q.batch_size = len(reqs)
                c += q.r
            ev = torch.cuda.Event()
            ev.record(s)
        return _Batch(slot, e, reqs, ev, t_now)

    def serve(self, requests, realtime: bool = True) -> ServeStats:
        """Serve `requests` (sorted by t_arrive).  realtime: a request becomes visible at
        its arrival time; otherwise all are visible at once (saturation throughput)."""
        torch.cuda.synchronize(self.cache.device)
        reqs = sorted(requests, key=lambda q: q.t_arrive)
        queues: OrderedDict[ProblemSpec, deque] = OrderedDict()
        free = list(range(len(self.streams)))
        inflight: list[_Batch] = []
        i = done = nb = 0
        n = len(reqs)
        t0 = time.perf_counter()
        if not realtime:
            for q in reqs:
                q.t_arrive = 0.0
        while done < n:
            now = time.perf_counter() - t0
            while i < n and reqs[i].t_arrive <= now:
                queues.setdefault(reqs[i].spec, deque()).append(reqs[i])
                i += 1
            still = []
            for b in inflight:
                if b.event.query():
                    t = time.perf_counter() - t0
                    for q in b.reqs:
                        q.t_done = t
                    done += len(b.reqs)
                    b.entry.busy = False
                    self.cache.release(b.entry)
                    free.append(b.slot)
                else:
                    still.append(b)
            inflight = still
            while free:
                heads = [(dq[0].t_arrive, spec) for spec, dq in queues.items()
                         if dq and not (spec in self.cache.ops and self.cache.ops[spec].busy)]
                if not heads:
                    break
                launched = False
                for _, spec in sorted(heads, key=lambda h: h[0]):
                    e = self.cache.acquire(spec)
                    if e is None:
                        continue  # not admissible now; try the next operator
                    dq = queues[spec]
                    batch, cols = [], 0
                    while dq and (not batch or cols + dq[0].r <= self.max_cols):
                        q = dq.popleft()
                        batch.append(q)
                        cols += q.r
                    inflight.append(self._launch(free.pop(), e, batch, now))
                    nb += 1
                    launched = True
                    break
                if not launched:
                    if not inflight:  # nothing will free budget: the operator cannot fit
                        raise RuntimeError("an operator does not fit the HBM budget of the "
                                           "factor cache")
                    break
        wall = max(q.t_done for q in reqs) - min(q.t_arrive for q in reqs)
        st = _stats(reqs, wall, nb)
        st.extra = {"cache_hits": self.cache.hits, "cache_misses": self.cache.misses,
                    "evictions": self.cache.evictions, "cache_bytes": self.cache.nbytes}
        return st


def serve_plain(cache: FactorCache, requests, realtime: bool = True) -> ServeStats:
    """The plain FP64 GPU path: one request at a time in arrival order, eager launches on
    the current stream, completion awaited before the next request starts."""
    torch.cuda.synchronize(cache.device)
    reqs = sorted(requests, key=lambda q: q.t_arrive)
    if not realtime:
        for q in reqs:
            q.t_arrive = 0.0
    t0 = time.perf_counter()
    for q in reqs:
        while time.perf_counter() - t0 < q.t_arrive:
            pass
        e = cache.acquire(q.spec)
        if e is None:
            raise RuntimeError("HBM budget too small for one operator")
        rhs = e.op.rhs_from_boundary(q.F)
        q.X = e.fac.solve(rhs)
        torch.cuda.current_stream().synchronize()
        q.t_done = time.perf_counter() - t0
        q.batch_size = 1
        cache.release(e)
    wall = max(q.t_done for q in reqs) - min(q.t_arrive for q in reqs)
    return _stats(reqs, wall, len(reqs))
