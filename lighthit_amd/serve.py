"""Serving loop of the shower route on one GPU.

Events are submitted as they arrive; a pool of worker threads, each with its
own HIP stream (and so its own workspaces, which are kept per stream), runs
``ShowerTransport`` calls on the device concurrently: one event's host steps
overlap another's kernels, and kernels of different events share the GPU.
An event is admitted when a worker is free and its estimated device-memory
footprint fits the budget beside the events already running. Every event's
result is the same as when it runs alone: the kernels of one call read and
write only that call's buffers and sum in a fixed order.

``ProcessShowerServer`` runs the same loop over worker processes, each with its
own route on the same GPU: the host steps of different events then run in
parallel rather than under one interpreter lock, and their kernels still share
the device.

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
    """Prior estimate of one event's device bytes before any event has been
    measured: the source upload and compile, the order-1 pencils and (pencil,
    OM) sort for up to ``oms`` OMs, and the order >=2 spectra; a fixed 256 MiB
    covers the per-stream workspaces."""
    elements = getattr(source, "elements", source)
    n = len(elements)
    return int(256 * 2 ** 20 + 2048 * n + 64 * n * min(oms, 512))


class FootprintModel:
    """Device bytes an event needs, learned from the events served so far.

    Each worker process reports the peak device memory it allocated during an
    event (or a batch of events: the caching allocator's high-water mark above
    its level before the call). An event's need is a fixed part (workspaces
    sized by the tables and the OM count) plus a part that grows with its
    elements: the model keeps the largest peak of small events (below
    ``SMALL`` elements) and, from the larger ones, the largest excess per
    element above the smallest peak seen; it admits an event of ``n``
    elements as ``margin * max(small peak, smallest peak + per_element * n)``.
    Small events alone would overstate the per-element part (their peaks are
    mostly the fixed part), so only large ones set it. Before the first
    report it uses :func:`footprint_bytes`.
    """

    SMALL = 16384

    def __init__(self, oms, margin=1.25):
        self.oms = int(oms)
        self.margin = float(margin)
        self.fixed = None
        self.small_peak = 0.0
        self.per_element = 0.0
        self.samples = 0
        self._lock = threading.Lock()

    def need(self, elements):
        n = len(elements)
        with self._lock:
            if self.fixed is None:
                return footprint_bytes(elements, self.oms)
            return int(self.margin * max(self.small_peak, self.fixed + self.per_element * n))

    def record(self, elements, peak_bytes):
        n = max(len(elements), 1)
        with self._lock:
            self.samples += 1
            if self.fixed is None or peak_bytes < self.fixed:
                self.fixed = float(peak_bytes)
            if n < self.SMALL:
                self.small_peak = max(self.small_peak, float(peak_bytes))
            else:
                self.per_element = max(self.per_element, (peak_bytes - self.fixed) / n)


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


# ------------------------------------------------------------ worker processes

def _pack_response(response):
    """A response as plain arrays, with only the bin rows that hold values (the
    full (OMs, bins, 3) array is mostly zero rows)."""
    components = response.components_pe
    rows = np.flatnonzero(components.reshape(len(components), -1).any(axis=1))
    return {"omega": response.omega_per_ns, "edges": response.relative_time_edges_ns,
            "spectrum": response.spectrum_ge2_pe, "rows": rows,
            "row_components": np.ascontiguousarray(components[rows]),
            "components_shape": components.shape,
            "charge": response.charge_components_pe, "origins": response.time_origin_ns,
            "active": response.active, "method": response.method,
            "metadata": response.metadata, "orders": tuple(response.computed_orders)}


def _unpack_response(packed, detector):
    from .route import HybridTransportResponse
    components = np.zeros(packed["components_shape"])
    components[packed["rows"]] = packed["row_components"]
    return HybridTransportResponse(detector, packed["omega"], packed["edges"], packed["spectrum"],
                                   components, packed["charge"], packed["origins"],
                                   packed["active"], packed["method"], packed["metadata"],
                                   packed["orders"])


class StagedSource:
    """Handle of a source staged in every worker process (its elements already
    in device memory there); ``ProcessShowerServer.stage`` makes it."""

    def __init__(self, key, count):
        self.key = key
        self.count = count

    def __len__(self):
        return self.count


def _process_main(factory, device, threads, tasks, results, slot, memory_fraction=None,
                  graphs=False):
    """One worker process: build the route, then answer tasks until ``None``.
    ``device`` None leaves the GPU alone (host-only routes). The process uses at
    most ``memory_fraction`` of the device memory; an event that runs out of it
    is run once more after the cached workspaces are released."""
    from threadpoolctl import threadpool_limits
    limits = threadpool_limits(limits=threads)
    drain = None
    out_of_memory = ()
    try:
        if device is not None:
            import torch
            torch.cuda.set_device(torch.device(device))
            if memory_fraction is not None:
                torch.cuda.set_per_process_memory_fraction(float(memory_fraction),
                                                           torch.device(device))
            drain = torch.cuda.synchronize
            out_of_memory = (torch.cuda.OutOfMemoryError,)
        route = factory()
        route.tables()
        route.order0()
        results.put(("ready", slot, None, None))
    except BaseException as error:      # reported to the parent, which raises it
        results.put(("failed", slot, None, f"{type(error).__name__}: {error}"))
        return
    staged = {}
    while True:
        item = tasks.get()
        if item is None:
            break
        kind, key, payload, call_kwargs = item
        if kind == "stage":
            try:
                from .route import ResidentShower
                staged[key] = ResidentShower(payload, device) if device is not None else payload
                results.put(("staged", slot, key, None))
            except BaseException as error:
                results.put(("stage_error", slot, key, f"{type(error).__name__}: {error}"))
            continue
        if kind == "release":
            staged.pop(key, None)
            continue
        # a task is a batch of events (keys, payloads): one event runs as a plain
        # call, several through route.batch (one fused order-0 launch for all)
        keys, payloads = (key, payload) if kind == "batch" else ([key], [payload])
        sent = set()
        try:
            sources = [staged[p.key] if isinstance(p, StagedSource) else p for p in payloads]
            retries = 0
            base = 0
            if device is not None:
                torch.cuda.reset_peak_memory_stats()
                base = torch.cuda.memory_allocated()
            try:
                if len(sources) == 1:
                    responses = [route(sources[0], **call_kwargs)]
                else:
                    responses = route.batch(sources, graph=graphs, **call_kwargs)
                    if any(isinstance(r, out_of_memory) for r in responses):
                        raise next(r for r in responses if isinstance(r, out_of_memory))
            except out_of_memory:
                # counted and reported: the admission model underestimated the
                # events; each runs once more on its own
                from . import _lib
                retries = 1
                _lib.release_workspaces()
                torch.cuda.empty_cache()
                responses = []
                for source in sources:
                    try:
                        responses.append(route(source, **call_kwargs))
                    except (ValueError, RuntimeError, TypeError) as error:
                        responses.append(error)
            if drain is not None:
                drain()
            peak = (torch.cuda.max_memory_allocated() - base) if device is not None else 0
            for k, response in zip(keys, responses):
                if isinstance(response, BaseException):
                    results.put(("error", slot, k, (type(response).__name__, str(response))))
                else:
                    response.metadata["out_of_memory_retries"] = retries
                    response.metadata["device_peak_bytes"] = int(peak)
                    response.metadata["batch_events"] = len(sources)
                    response.metadata["batch_elements"] = int(sum(len(getattr(x, "elements", x))
                                                                  for x in sources))
                    results.put(("done", slot, k, _pack_response(response)))
                sent.add(k)
        except BaseException as error:
            for k in keys:
                if k not in sent:
                    results.put(("error", slot, k, (type(error).__name__, str(error))))
        results.put(("free", slot, None, None))
    limits.restore_original_limits()


class ProcessShowerServer:
    """``ShowerTransport`` calls on one GPU from ``processes`` worker processes.

    ``factory`` is a picklable callable that returns the route; each process
    calls it once, so every process holds the tables resident in HBM and uses
    at most ``memory_fraction`` of the device memory (default: an equal share
    of 90% among the workers and the caller). An event is sent to a free
    process when its footprint (:class:`FootprintModel`, learned from the
    workers' measured peaks) fits the budget (the workers' shares together)
    beside the events already running; the response comes back as arrays.
    ``stats`` counts served events, out-of-memory retries (an event that ran
    out of its worker's share runs once more after the worker drops its
    cached workspaces; the response's ``metadata["out_of_memory_retries"]``
    says so) and the largest peak. ``submit`` returns a ``Future`` as ``ShowerServer.submit`` does.
    Sources travel as host element arrays (a ``ResidentShower`` is sent as its
    host elements), or are staged ahead in device memory with ``stage``.

    Continuous batching (``batch`` > 1): when a process is free, the event at
    the head of the queue and the events already waiting behind it (at most
    ``batch``, while they fit the memory budget, and only events without their
    own call arguments) go to it as one task, which it runs through
    ``route.batch`` (their order-0 charges in one fused persistent launch;
    with ``graphs`` that launch is replayed from a HIP graph captured once per
    capacity); events arriving meanwhile wait for the next free process. Each
    response is the plain call's.
    """

    def __init__(self, factory, *, processes=2, detector, memory_budget_bytes=None,
                 device="cuda:0", threads_per_process=None, memory_fraction=None,
                 start_timeout_s=600.0, batch=1, graphs=False):
        import multiprocessing as mp
        import os
        if processes < 1:
            raise ValueError("processes must be positive")
        if batch < 1:
            raise ValueError("batch must be positive")
        self.detector = detector
        self.device = device
        self.batch = int(batch)
        # each worker's share of the device memory (the caller keeps one share)
        fraction = memory_fraction if memory_fraction is not None else 0.9 / (processes + 1)
        if memory_budget_bytes is None:
            if device is None:
                raise ValueError("a host-only server needs memory_budget_bytes")
            import torch
            _, total = torch.cuda.mem_get_info(torch.device(device))
            # every worker holds at most its share: events in flight together
            # must fit the workers' shares together
            memory_budget_bytes = int(fraction * total * processes)
        self.budget = int(memory_budget_bytes)
        self.footprint = FootprintModel(len(detector))
        self.stats = {"events": 0, "out_of_memory_retries": 0, "largest_peak_bytes": 0}
        share = int(os.environ.get("OMP_NUM_THREADS", 0)) or (os.cpu_count() or 1)
        threads = threads_per_process or max(1, share // (processes + 1))
        context = mp.get_context("spawn")
        self._results = context.Queue()
        self._tasks = [context.Queue() for _ in range(processes)]
        self._processes = [context.Process(target=_process_main,
                                           args=(factory, device, threads, self._tasks[i],
                                                 self._results, i, fraction, graphs),
                                           name=f"shower-process-{i}", daemon=True)
                           for i in range(processes)]
        for process in self._processes:
            process.start()
        self._lock = threading.Condition()
        self._idle = []
        self._inflight = 0
        self._pending = {}
        self._staging = {}
        self._waiting = queue.Queue()
        self._closed = False
        self._next_key = 0
        started = time.perf_counter()
        ready = 0
        while ready < processes:
            remaining = start_timeout_s - (time.perf_counter() - started)
            try:
                kind, slot, _, message = self._results.get(timeout=min(max(remaining, 0.1), 1.0))
            except queue.Empty:
                dead = [i for i, p in enumerate(self._processes) if not p.is_alive()]
                if dead or remaining <= 0:
                    codes = {i: self._processes[i].exitcode for i in dead}
                    self._terminate()
                    raise RuntimeError("worker processes did not start: "
                                       + (f"exited {codes}" if dead else "not in time")) from None
                continue
            if kind == "failed":
                self._terminate()
                raise RuntimeError(f"worker process {slot} failed to start: {message}")
            self._idle.append(slot)
            ready += 1
        self._collector = threading.Thread(target=self._collect, name="shower-collector",
                                           daemon=True)
        self._dispatcher = threading.Thread(target=self._dispatch, name="shower-dispatcher",
                                            daemon=True)
        self._collector.start()
        self._dispatcher.start()

    def submit(self, source, **call_kwargs):
        from .route import _elements_of
        if self._closed:
            raise RuntimeError("the server is closed")
        future = Future()
        future.submitted = time.perf_counter()
        elements = source if isinstance(source, StagedSource) else _elements_of(source)
        self._waiting.put((future, elements, call_kwargs, None))
        return future

    def stage(self, source):
        """Upload ``source`` to the device in every worker process ahead of its
        call; returns the handle to ``submit``. Each staged handle is a separate
        source (the per-event compile runs for it)."""
        from .route import _elements_of
        if self._closed:
            raise RuntimeError("the server is closed")
        elements = _elements_of(source)
        with self._lock:
            key = self._next_key
            self._next_key += 1
            self._staging[key] = []
        for tasks in self._tasks:
            tasks.put(("stage", key, elements, None))
        with self._lock:
            while len(self._staging[key]) < len(self._processes):
                if not all(process.is_alive() for process in self._processes):
                    raise RuntimeError("a worker process ended while staging")
                self._lock.wait(timeout=1.0)
            errors = [m for m in self._staging.pop(key) if m is not None]
        if errors:
            raise RuntimeError(f"staging failed: {errors[0]}")
        return StagedSource(key, len(elements))

    def release(self, handle):
        """Free a staged source in every worker process."""
        for tasks in self._tasks:
            tasks.put(("release", handle.key, None, None))

    def map(self, sources, **call_kwargs):
        futures = [self.submit(source, **call_kwargs) for source in sources]
        return [f.result() for f in futures]

    def _dispatch(self):
        held = None
        while True:
            item = held if held is not None else self._waiting.get()
            held = None
            if item is None:
                return
            # the model at dispatch time (it learns from the events finished so far)
            group = [item + (self.footprint.need(item[1]),)]
            with self._lock:
                self._idle = [i for i in self._idle if self._processes[i].is_alive()]
                while not self._idle or (self._inflight > 0
                                         and self._inflight + group[0][4] > self.budget):
                    if not any(process.is_alive() for process in self._processes):
                        break
                    self._lock.wait(timeout=1.0)
                    self._idle = [i for i in self._idle if self._processes[i].is_alive()]
                if not self._idle:
                    future = item[0]
                    future.latency_s = time.perf_counter() - future.submitted
                    future.set_exception(RuntimeError("no worker process is running"))
                    continue
                # continuous batching: events already waiting join this task
                # while they fit the memory budget and the batch size
                need = group[0][4]
                while len(group) < self.batch:
                    try:
                        extra = self._waiting.get_nowait()
                    except queue.Empty:
                        break
                    if extra is None:
                        held = None
                        self._waiting.put(None)
                        break
                    extra_need = self.footprint.need(extra[1])
                    if (extra[2] or item[2]       # events with their own arguments run alone
                            or self._inflight + need + extra_need > self.budget):
                        held = extra
                        break
                    group.append(extra + (extra_need,))
                    need += extra_need
                slot = self._idle.pop(0)
                self._inflight += need
                keys = []
                for future, elements, call_kwargs, _, event_need in group:
                    key = self._next_key
                    self._next_key += 1
                    self._pending[key] = (future, event_need, slot, len(elements))
                    keys.append(key)
            if len(group) == 1:
                self._tasks[slot].put(("run", keys[0], group[0][1], group[0][2]))
            else:
                self._tasks[slot].put(("batch", keys, [g[1] for g in group], group[0][2]))

    def _collect(self):
        errors = {"ValueError": ValueError, "RuntimeError": RuntimeError,
                  "TypeError": TypeError, "MemoryError": MemoryError}
        while True:
            try:
                kind, slot, key, payload = self._results.get(timeout=1.0)
            except queue.Empty:
                self._fail_dead()
                continue
            if kind == "stop":
                return
            if kind in ("staged", "stage_error"):
                with self._lock:
                    self._staging[key].append(payload)
                    self._lock.notify_all()
                continue
            if kind == "free":
                with self._lock:
                    self._idle.append(slot)
                    self._lock.notify_all()
                continue
            with self._lock:
                future, need, _, count = self._pending.pop(key)
                self._inflight -= need
                self._lock.notify_all()
            future.latency_s = time.perf_counter() - future.submitted
            if kind == "done":
                meta = payload["metadata"]
                peak = int(meta.get("device_peak_bytes", 0))
                if peak > 0:
                    # a batch's peak belongs to all of its elements together
                    self.footprint.record(range(int(meta.get("batch_elements", count))), peak)
                with self._lock:
                    self.stats["events"] += 1
                    self.stats["out_of_memory_retries"] += int(meta.get("out_of_memory_retries", 0))
                    self.stats["largest_peak_bytes"] = max(self.stats["largest_peak_bytes"], peak)
                future.set_result(_unpack_response(payload, self.detector))
            else:
                name, message = payload
                error = errors.get(name)
                future.set_exception(error(message) if error is not None
                                     else RuntimeError(f"{name}: {message}"))

    def _fail_dead(self):
        """Events held by a worker process that has ended fail with an error;
        the process is not replaced."""
        with self._lock:
            dead = {i for i, process in enumerate(self._processes) if not process.is_alive()}
            lost = [key for key, (_, _, slot, _) in self._pending.items() if slot in dead]
            for key in lost:
                future, need, slot, _ = self._pending.pop(key)
                self._inflight -= need
                future.latency_s = time.perf_counter() - future.submitted
                future.set_exception(RuntimeError(f"worker process {slot} ended "
                                                  f"(exit code {self._processes[slot].exitcode})"))
            if lost:
                self._lock.notify_all()

    def _terminate(self):
        for process in self._processes:
            if process.is_alive():
                process.terminate()
        for process in self._processes:
            process.join(timeout=30)

    def close(self):
        if self._closed:
            return
        self._closed = True
        self._waiting.put(None)
        self._dispatcher.join()
        with self._lock:
            while self._pending:
                self._lock.wait()
        for tasks in self._tasks:
            tasks.put(None)
        for process in self._processes:
            process.join(timeout=60)
        self._results.put(("stop", None, None, None))
        self._collector.join()
        self._terminate()

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
