"""Fixed-shape solves captured once as a HIP graph and replayed.

Solves built from many small launches -- the HBS Thomas solve issues thousands of
GEMMs, ULV steps and gathers at the 3-D interface size -- spend most of their time in
host-side launch overhead.  Their shapes are fixed once the factorization exists, so
the whole launch sequence is recorded once (torch.cuda.CUDAGraph around the library's
stream-ordered calls; the library clears memory with fill kernels, never memsets) and
replayed: a call copies the right-hand sides into the captured input panel, replays,
and returns the captured output panel (overwritten by the next call).
"""
from __future__ import annotations

import gc

import torch


class CapturedSolve:
    """solve(B) -> X for B of one fixed shape, dtype and device, replayed from a graph."""

    def __init__(self, solve, shape, dtype=torch.float64, device=None):
        dev = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.B = torch.zeros(shape, dtype=dtype, device=dev)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            solve(self.B)  # warm-up on the capture stream (code objects, allocator pools)
        torch.cuda.current_stream(dev).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        # no garbage collection while capturing: cyclic garbage freed then (another graph,
        # pinned memory) would touch the device or other streams inside the capture
        gc_was = gc.isenabled()
        gc.disable()
        try:
            with torch.cuda.graph(self.graph):
                self.X = torch.as_tensor(solve(self.B), device=dev)
        finally:
            if gc_was:
                gc.enable()

    def __call__(self, B: torch.Tensor) -> torch.Tensor:
        if B.shape != self.B.shape:
            raise ValueError(f"CapturedSolve: shape {tuple(B.shape)} != captured "
                             f"{tuple(self.B.shape)}")
        self.B.copy_(B)
        self.graph.replay()
        return self.X
