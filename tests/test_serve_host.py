"""Host logic of the serving loop (admission, ordering, errors) with a stand-in
route and no device: the worker context is replaced by a no-op."""
import contextlib
import threading
import time

import numpy as np
import pytest

from lighthit_amd import serve


class _Kernel:
    def __init__(self, n):
        self.detector = np.zeros((n, 3))


class _Route:
    """Records concurrency; returns the source's value doubled."""

    def __init__(self):
        self.kernel = _Kernel(4)
        self.lock = threading.Lock()
        self.running = 0
        self.peak = 0

    def tables(self):
        return None

    def order0(self):
        return None

    def __call__(self, source, **kwargs):
        with self.lock:
            self.running += 1
            self.peak = max(self.peak, self.running)
        time.sleep(0.01)
        with self.lock:
            self.running -= 1
        if source.value < 0:
            raise ValueError("negative source")
        return 2 * source.value + kwargs.get("offset", 0)


class _Source:
    def __init__(self, value, n=10):
        self.value = value
        self.elements = np.zeros(n)


class _HostServer(serve.ShowerServer):
    def _stream_context(self):
        return contextlib.nullcontext(), lambda: None


def test_results_in_order_and_concurrent():
    route = _Route()
    with _HostServer(route, workers=3, memory_budget_bytes=10 ** 12) as server:
        out = server.map([_Source(i) for i in range(12)], offset=1)
    assert out == [2 * i + 1 for i in range(12)]
    assert route.peak >= 2


def test_budget_serialises_large_events():
    route = _Route()
    need = serve.footprint_bytes(_Source(1), 4)
    with _HostServer(route, workers=3, memory_budget_bytes=need) as server:
        out = server.map([_Source(i) for i in range(6)])
    assert out == [2 * i for i in range(6)]
    assert route.peak == 1


def test_errors_reach_the_caller_and_latency_is_set():
    route = _Route()
    with _HostServer(route, workers=2, memory_budget_bytes=10 ** 12) as server:
        good, bad = server.submit(_Source(3)), server.submit(_Source(-1))
        assert good.result() == 6
        with pytest.raises(ValueError, match="negative"):
            bad.result()
    assert good.latency_s > 0 and bad.latency_s > 0
    summary = serve.latency_summary([good, bad])
    assert summary["p50"] <= summary["p99"] <= summary["max"]


def test_closed_server_refuses_work():
    server = _HostServer(_Route(), workers=1, memory_budget_bytes=10 ** 12)
    server.close()
    with pytest.raises(RuntimeError, match="closed"):
        server.submit(_Source(1))


def test_footprint_grows_with_elements():
    assert serve.footprint_bytes(_Source(0, 1000), 48) < serve.footprint_bytes(
        _Source(0, 100000), 48)
