"""hipGraph capture of a fixed operator action.

A Krylov recurrence applies one effective Hamiltonian (environments and MPO tensor fixed) many
times.  ``CapturedApply`` records that action's launches once, as a graph on a private stream,
and replays it: one launch call per application instead of one host round trip per
grouped-GEMM stage, so the stages run back to back with no host gaps.  While the graph lives the
library keeps the plan images and scratch buffers it reads alive (``bug_graph_capture_begin`` /
``bug_graph_release`` in include/bug.h).  The reference has no counterpart: its
``expv``/``lanczos_expv`` (src/BUG/Core/LocalSolvers/krylov.jl:128-442) call the action eagerly.
"""
from typing import Callable

import torch

from . import _lib
from .blocks import BlockTensor


class CapturedApply:
    """``action(x)`` captured for one input layout: ``cap(v)`` copies ``v`` into the graph's
    input, replays the graph on the current stream and returns the graph's output tensor, which
    the next replay overwrites (clone it to keep it).

    ``action`` must launch only library entry points on the current stream (e.g.
    ``lambda x: environments.apply_one_site(L, W, R, x)``).  It runs once eagerly on the private
    stream first, so that its plans, schedules and scratch exist before capture; a capture that
    would need any of them new fails with ``BugError``.
    """

    def __init__(self, action: Callable[[BlockTensor], BlockTensor], x: BlockTensor):
        self._ctx = _lib.context()
        self._stream = torch.cuda.Stream()
        self._pinned = False
        self.x = BlockTensor(x.layout, x.data.clone())
        self._stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(self._stream):
            action(self.x)
        self._stream.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        lib, h = self._ctx._lib, self._ctx.handle
        try:
            with torch.cuda.graph(self.graph, stream=self._stream,
                                  capture_error_mode="thread_local"):
                self._ctx.check(lib.bug_graph_capture_begin(h, self._stream.cuda_stream))
                self._pinned = True
                try:
                    self.y = action(self.x)
                finally:
                    self._ctx.check(lib.bug_graph_capture_end(h))
        except BaseException:
            self.close()
            raise

    def __call__(self, v: BlockTensor) -> BlockTensor:
        if v.layout != self.x.layout:
            raise ValueError("input layout differs from the captured one")
        self.x.data.copy_(v.data)
        self.graph.replay()
        return self.y

    def replay(self) -> BlockTensor:
        """Replay on the current input (``self.x``)."""
        self.graph.replay()
        return self.y

    def close(self):
        """Drop the graph and let the library free what it kept for it."""
        if self._pinned:
            self.graph = None
            self._pinned = False
            self._ctx.check(self._ctx._lib.bug_graph_release(self._ctx.handle,
                                                              self._stream.cuda_stream))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
