"""Captured HIP graphs of fixed launch sequences (a sampler block, a fragment batch's
block, a resampling tail).

A ``Replay`` wraps a callable that enqueues library launches on torch's current stream
with fixed buffers.  The first call runs it eagerly (kernel attributes, lazily sized
workspace and torch's allocator settle); the second call captures it into a graph
(torch.cuda.CUDAGraph, thread-local capture mode so that other host threads may keep
allocating) and replays it; later calls replay.  Capture records kernel arguments by
value, so the callable must only touch buffers that outlive the Replay and must not
read anything back to the host; every random draw it makes has to come from a
device-resident key.

During a capture the Python garbage collector is paused: a finalizer that ran then
(a context's hipFree, another graph's destruction) would be an operation the capturing
thread may not perform.  A bound-method callable is held weakly, so an owner that keeps
its Replay does not form a reference cycle that only the collector could break.
"""

import gc
import types
import weakref

import torch


class Replay:
    def __init__(self, fn, use_graph=True):
        self._fn = weakref.WeakMethod(fn) if isinstance(fn, types.MethodType) else (lambda: fn)
        self.use_graph = use_graph
        self.graph = None
        self.warm = False
        self._side = None
        self._stream = None  # the stream the graph was captured on

    @property
    def fn(self):
        f = self._fn()
        if f is None:
            raise RuntimeError("the object that owns this captured block no longer exists")
        return f

    def __call__(self):
        if not self.use_graph:
            self.fn()
            return
        if not self.warm:
            self.fn()
            self.warm = True
            return
        if self.graph is None:
            # capture on the caller's stream (the library objects bound to it keep their
            # stream, so nothing in the capture changes or synchronises streams); the
            # default stream cannot capture, so a side stream stands in for it
            cur = torch.cuda.current_stream()
            if cur.cuda_stream == 0:
                if self._side is None:
                    self._side = torch.cuda.Stream()  # default priority
                cap = self._side
                cap.wait_stream(cur)
            else:
                cap = cur
            fn = self.fn
            g = torch.cuda.CUDAGraph()
            gc_on = gc.isenabled()
            gc.disable()
            try:
                with torch.cuda.graph(g, stream=cap, capture_error_mode="thread_local"):
                    fn()
            finally:
                if gc_on:
                    gc.enable()
            if cap is not cur:
                cur.wait_stream(cap)
            self.graph = g
            self._stream = cap
        self.graph.replay()

    def release(self):
        """Drop the captured graph once the GPU has finished every replay of it."""
        if self.graph is not None:
            torch.cuda.synchronize()
        self.graph = None
        self.warm = False
