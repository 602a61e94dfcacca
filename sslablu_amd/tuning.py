"""GEMM tile-policy autotuner.

Every GEMM launch of libsslab_hip.so picks its tile shape and K split from a
per-thread policy (``sslab_gemm_set_policy``): the workgroup count a tile must reach
before a smaller one is tried, the split-K workgroup target, the tile ids for
mid-size and split launches and a cap on the largest tile.  The built-in defaults
are a compromise over problem shapes; this module measures the candidates on the
caller's own operation (a block-tridiagonal solve of a given size, say) and keeps the
fastest policy per key, in memory and optionally in a JSON file named by
``SSLAB_TUNE_CACHE``.

Results never depend on the policy beyond floating-point summation order (split-K
changes the order of the K sum), so a tuned policy is a pure performance choice.
"""
from __future__ import annotations

import ctypes
import itertools
import json
import os
import threading
from contextlib import contextmanager
from dataclasses import astuple, dataclass

import torch

from . import _lib
from ._lib import call


@dataclass(frozen=True)
class Policy:
    target_wgs: int = 1024
    split_wgs: int = 512
    max_ksplit: int = 64
    min_split_k: int = 256
    cfg_mid: int = 3
    cfg_split: int = 3
    max_tile: int = 0


DEFAULT = Policy()

_cache: dict[str, Policy] = {}
_lock = threading.Lock()
_loaded = False


# The calling thread's policy and backward-level table as last set through this module
# (the library keeps them per thread too): equal settings skip the library call, so the
# per-call policy scopes of the solvers cost no host time once a policy is standing.
_tls = threading.local()


def set_policy(p: Policy) -> None:
    """Set the calling thread's GEMM policy in the library."""
    if getattr(_tls, "policy", None) == p:
        return
    arr = (ctypes.c_int * 7)(*astuple(p))
    try:
        call("sslab_gemm_set_policy", ctypes.addressof(arr), 7)
    except Exception:
        _tls.policy = None
        raise
    _tls.policy = p


def set_backward_policies(pols) -> None:
    """Per-step tile policies of the red-black backward sweep for the calling thread's
    later solves (step u runs the sweep's level L-1-u); an empty list clears them."""
    pols = tuple(pols or ())
    if getattr(_tls, "bwd", None) == pols:
        return
    arr = (ctypes.c_int * (7 * max(1, len(pols))))()
    for u, p in enumerate(pols):
        arr[7 * u:7 * u + 7] = list(astuple(p))
    try:
        call("sslab_rb_set_backward_policies", ctypes.addressof(arr), len(pols))
    except Exception:
        _tls.bwd = None
        raise
    _tls.bwd = pols


@contextmanager
def backward_policies(pols):
    """Run the body with per-step backward-sweep policies `pols` (None: leave them)."""
    if pols is None:
        yield
        return
    old = getattr(_tls, "bwd", None)
    set_backward_policies(pols)
    try:
        yield
    finally:
        set_backward_policies(old if old is not None else ())


def get_policy() -> Policy:
    p = getattr(_tls, "policy", None)
    if p is not None:
        return p
    arr = (ctypes.c_int * 7)()
    call("sslab_gemm_get_policy", ctypes.addressof(arr), 7)
    _tls.policy = Policy(*arr)
    return _tls.policy


@contextmanager
def policy(p: Policy | None):
    """Run the body under policy `p` (None: leave the current policy)."""
    if p is None:
        yield
        return
    old = get_policy()
    set_policy(p)
    try:
        yield
    finally:
        set_policy(old)


def candidates() -> list[Policy]:
    """The search space: workgroup targets, split-K targets and tile ids (split launches
    of few columns favour the small 64x32 / 32x32 tiles with many K parts: the
    workgroups of a narrow launch are latency-bound, so more of them in flight help)."""
    out = []
    for target, split, cmid, csplit, mt in itertools.product(
            (512, 1024), (512, 1024, 2048, 4096), (2, 3, 10), (2, 3, 4, 5, 6, 9, 10, 12, 13, 14),
            (0, 1)):
        out.append(Policy(target, split, 64, 256, cmid, csplit, mt))
    return out


def _cache_path():
    return os.environ.get("SSLAB_TUNE_CACHE")


def _load():
    global _loaded
    if _loaded:
        return
    _loaded = True
    path = _cache_path()
    if path and os.path.exists(path):
        with open(path) as f:
            for k, v in json.load(f).items():
                _cache[k] = Policy(*v)


def _save():
    path = _cache_path()
    if path:
        with open(path, "w") as f:
            json.dump({k: list(astuple(v)) for k, v in _cache.items()}, f, indent=1)


def remember(key: str, p: Policy) -> None:
    """Record `p` as the policy for `key` (as autotune does), e.g. a policy measured
    elsewhere; persisted when SSLAB_TUNE_CACHE is set."""
    with _lock:
        _load()
        _cache[key] = p
        _save()


def lookup(key: str) -> Policy | None:
    with _lock:
        _load()
        return _cache.get(key)


def autotune(fn, key: str, cands=None, reps: int = 10, stream=None) -> tuple[Policy, float]:
    """Time `fn()` (GPU work on the current stream) under every candidate policy and
    remember the fastest under `key`.  Returns (policy, microseconds per call)."""
    _lib.require_gpu()
    s = stream if stream is not None else torch.cuda.current_stream()
    cands = list(cands) if cands is not None else candidates()
    if DEFAULT not in cands:
        cands.append(DEFAULT)
    best = None
    old = get_policy()
    try:
        for p in cands:
            set_policy(p)
            fn()  # warm-up (code objects, caches)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(reps):
                fn()
            e1.record(s)
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) * 1e3 / reps
            if best is None or t < best[1]:
                best = (p, t)
    finally:
        set_policy(old)
    with _lock:
        _load()
        _cache[key] = best[0]
        _save()
    return best
