"""Serving loop of the shower route on one GPU.

Events are submitted as they arrive; a pool of worker threads, each with its
own HIP stream (and so its own workspaces, which are kept per stream), runs
``ShowerTransport`` calls on the device concurrently: one event's host steps
overlap another's kernels, and kernels of different events share the GPU.
An event is admitted when a worker is free and its estimated device-memory
footprint fits the budget beside the events already running. Every event's
result is the same as when it runs alone: the kernels of one call read and
write only that call's buffers and sum in a fixed order.

The reference evaluates one event per call (``shower/api.py:176-190``); this
loop is the serving form around that call.
"""
from __future__ import annotations

import queue
import threading
import time
from concurrent.futures import Future

import numpy as np


def footprint_bytes(source, oms):
    """Conservative device bytes of one event: the source upload and compile,
    the order-1 pencils and (pencil, OM) sort for up to ``oms`` OMs, and the
    order >=2 spectra; a fixed 256 MiB covers the per-stream workspaces."""
    elements = getattr(source, "elements", source)
    n = len(elements)
    return int(256 * 2 ** 20 + 2048 * n + 64 * n * min(oms, 512))


class ShowerServer:
    """Concurrent ``ShowerTransport`` calls on one GPU.

    ``submit(source, **call_kwargs)`` returns a ``concurrent.futures.Future``
    of the response; its ``latency_s`` attribute holds the time from
    submission to completion once it is done.
    """

    def __init__(self, route, *, workers=2, memory_budget_bytes=None, device="cuda"):
        if workers < 1:
            raise ValueError("workers must be positive")
        self.route = route
        self.device = device
        if memory_budget_bytes is None:
            import torch
            free, _ = torch.cuda.mem_get_info(torch.device(device))
            memory_budget_bytes = int(0.5 * free)
        self.budget = int(memory_budget_bytes)
        self._queue = queue.Queue()
        self._admit = threading.Condition()
        self._inflight = 0
        self._closed = False
        # tables and resident parts are built once, before any worker runs
        route.tables()
        route.order0()
        self._threads = [threading.Thread(target=self._work, name=f"shower-worker-{i}",
                                          daemon=True) for i in range(workers)]
        for thread in self._threads:
            thread.start()

    def submit(self, source, **call_kwargs):
        if self._closed:
            raise RuntimeError("the server is closed")
        future = Future()
        future.submitted = time.perf_counter()
        need = footprint_bytes(source, len(self.route.kernel.detector))
        self._queue.put((future, source, call_kwargs, need))
        return future

    def map(self, sources, **call_kwargs):
        """Responses of ``sources`` in order (all submitted at once)."""
        futures = [self.submit(source, **call_kwargs) for source in sources]
        return [f.result() for f in futures]

    def _stream_context(self):
        """Context of one worker: its own HIP stream on the server's device."""
        import contextlib

        import torch
        device = torch.device(self.device)
        stream = torch.cuda.Stream(device)
        stack = contextlib.ExitStack()
        stack.enter_context(torch.cuda.device(device))
        stack.enter_context(torch.cuda.stream(stream))
        return stack, stream.synchronize

    def _work(self):
        context, drain = self._stream_context()
        with context:
            while True:
                item = self._queue.get()
                if item is None:
                    return
                future, source, call_kwargs, need = item
                with self._admit:
                    while self._inflight > 0 and self._inflight + need > self.budget:
                        self._admit.wait()
                    self._inflight += need
                try:
                    response = self.route(source, **call_kwargs)
                    drain()
                    future.latency_s = time.perf_counter() - future.submitted
                    future.set_result(response)
                except BaseException as error:   # the caller sees it on the future
                    future.latency_s = time.perf_counter() - future.submitted
                    future.set_exception(error)
                finally:
                    with self._admit:
                        self._inflight -= need
                        self._admit.notify_all()

    def close(self):
        if self._closed:
            return
        self._closed = True
        for _ in self._threads:
            self._queue.put(None)
        for thread in self._threads:
            thread.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def latency_summary(futures):
    """p50, p99 and max latency in ms of finished futures."""
    lat = np.asarray([f.latency_s for f in futures], float) * 1e3
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
            "max": float(lat.max())}
