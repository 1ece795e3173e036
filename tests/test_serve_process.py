"""Host logic of the process-based serving loop with a stand-in route in the
worker processes and no device: dispatch, responses carried back as arrays
(bin rows that hold values only), errors, and a factory that fails."""
import numpy as np
import pytest

from lighthit_amd import serve
from lighthit_amd.route import HybridTransportResponse

OMS, BINS = 6, 5


class _Elements:
    """Host elements of a stand-in event: its value and which OM it lights."""

    def __init__(self, value, om):
        self.coefficient0 = np.full(3, float(value))
        self.om = om

    def __len__(self):
        return len(self.coefficient0)


class _Route:
    def tables(self):
        return None

    def order0(self):
        return None

    def __call__(self, elements, *, offset=0.0):
        value = float(elements.coefficient0[0])
        if value < 0:
            raise ValueError("negative source")
        charge = np.zeros((OMS, 3))
        charge[elements.om] = [value, 2 * value, 3 * value + offset]
        components = np.zeros((OMS, BINS, 3))
        components[elements.om] = value
        spectrum = np.zeros((2, OMS), complex)
        spectrum[1, elements.om] = 1j * value
        active = charge.sum(axis=1) > 0
        return HybridTransportResponse(None, np.arange(2.0), np.arange(BINS + 1.0), spectrum,
                                       components, charge, np.full(OMS, 7.0), active, "shower",
                                       {"value": value})


    def batch(self, sources, graph=True, **call_kwargs):
        out = []
        for elements in sources:
            try:
                out.append(self(elements, **call_kwargs))
            except ValueError as error:
                out.append(error)
        return out


def _factory():
    return _Route()


def _failing_factory():
    raise OSError("tables missing")


def test_responses_round_trip_through_worker_processes():
    detector = np.zeros((OMS, 3))
    with serve.ProcessShowerServer(_factory, processes=2, detector=detector, device=None,
                                   memory_budget_bytes=10 ** 12,
                                   threads_per_process=1) as server:
        events = [_Elements(i + 1, i % OMS) for i in range(8)]
        futures = [server.submit(e, offset=0.5) for e in events]
        out = [f.result(timeout=120) for f in futures]
    for event, (i, r) in zip(events, enumerate(out)):
        want = _Route()(event, offset=0.5)
        assert r.detector is detector
        assert np.array_equal(r.components_pe, want.components_pe)
        assert np.array_equal(r.charge_components_pe, want.charge_components_pe)
        assert np.array_equal(r.spectrum_ge2_pe, want.spectrum_ge2_pe)
        assert np.array_equal(r.active, want.active)
        assert r.metadata["value"] == float(i + 1)
        assert r.metadata["out_of_memory_retries"] == 0
        assert futures[i].latency_s > 0
    assert server.stats["events"] == 8 and server.stats["out_of_memory_retries"] == 0


def test_worker_errors_reach_the_caller():
    with serve.ProcessShowerServer(_factory, processes=1, detector=np.zeros((OMS, 3)), device=None,
                                   memory_budget_bytes=10 ** 12,
                                   threads_per_process=1) as server:
        bad, good = server.submit(_Elements(-1, 0)), server.submit(_Elements(2, 1))
        with pytest.raises(ValueError, match="negative"):
            bad.result(timeout=120)
        assert good.result(timeout=120).charge_pe[1] == pytest.approx(12.0)


def test_failing_factory_raises_at_start():
    with pytest.raises(RuntimeError, match="tables missing"):
        serve.ProcessShowerServer(_failing_factory, processes=1, detector=np.zeros((OMS, 3)), device=None,
                                  memory_budget_bytes=10 ** 12, threads_per_process=1)


def test_packed_response_keeps_only_rows_with_values():
    response = _Route()(_Elements(3, 2))
    packed = serve._pack_response(response)
    assert list(packed["rows"]) == [2]
    back = serve._unpack_response(packed, None)
    assert np.array_equal(back.components_pe, response.components_pe)


def test_staged_sources_run_from_every_worker():
    with serve.ProcessShowerServer(_factory, processes=2, detector=np.zeros((OMS, 3)),
                                   device=None, memory_budget_bytes=10 ** 12,
                                   threads_per_process=1) as server:
        handles = [server.stage(_Elements(i + 1, i % OMS)) for i in range(4)]
        assert [len(h) for h in handles] == [3, 3, 3, 3]
        futures = [server.submit(h) for h in handles for _ in range(2)]
        out = [f.result(timeout=120) for f in futures]
        for h in handles:
            server.release(h)
    for j, r in enumerate(out):
        i = j // 2
        want = _Route()(_Elements(i + 1, i % OMS))
        assert np.array_equal(r.charge_components_pe, want.charge_components_pe)
        assert np.array_equal(r.components_pe, want.components_pe)


def test_footprint_model_learns_from_measured_peaks():
    model = serve.FootprintModel(oms=2304, margin=1.25)
    prior = model.need(range(1000))
    assert prior == serve.footprint_bytes(range(1000), 2304)
    model.record(range(1000), 300 * 2 ** 20)       # a small event: the fixed part
    model.record(range(100_000), 3300 * 2 ** 20)   # a large one: bytes per element
    per_element = (3300 - 300) * 2 ** 20 / 100_000
    assert model.need(range(50_000)) == int(1.25 * (300 * 2 ** 20 + per_element * 50_000))
    # a small event's peak is mostly the fixed part: it raises the floor of
    # the estimate, not the bytes per element
    model.record(range(10_000), 1000 * 2 ** 20)
    assert model.per_element == per_element
    assert model.need(range(2000)) == int(1.25 * 1000 * 2 ** 20)
    # a later, larger per-element peak of a large event raises the estimate
    model.record(range(20_000), 2000 * 2 ** 20)
    assert model.per_element > per_element


def test_waiting_events_are_batched_into_one_task():
    # one process, batches of up to 4: the events queued behind a running task
    # go to the process together; every result is the plain call's, and an
    # event that fails inside a batch fails alone
    with serve.ProcessShowerServer(_factory, processes=1, detector=np.zeros((OMS, 3)),
                                   device=None, memory_budget_bytes=10 ** 12,
                                   threads_per_process=1, batch=4) as server:
        events = [_Elements(i + 1, i % OMS) for i in range(9)] + [_Elements(-1, 0)]
        futures = [server.submit(e) for e in events]
        with pytest.raises(ValueError, match="negative"):
            futures[-1].result(timeout=120)
        out = [f.result(timeout=120) for f in futures[:-1]]
    for event, r in zip(events, out):
        want = _Route()(event)
        assert np.array_equal(r.charge_components_pe, want.charge_components_pe)
        assert np.array_equal(r.components_pe, want.components_pe)
    sizes = [r.metadata["batch_events"] for r in out]
    assert max(sizes) > 1 and max(sizes) <= 4
    assert server.stats["events"] == 9
