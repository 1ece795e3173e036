"""The whole route at the detector-scale event (2304 OMs, 50,000 elements, 8
wavelength nodes, 804 bins of 1 ns) against the oracle route's committed
fixture (tests/golden/make_detector_scale_route.py), at the bars of the plain
form: charge per order within 1e-5 relative (order 1 1e-6) on every OM above
threshold, the same ``active`` set except OMs within that tolerance of the
threshold, and bins per order against each active OM's bin L1. The precision
form ("mixed") must hold every active OM's charge within 1e-3 of the FP64 oracle.
Both order >=2 packs run: the seeded synthetic one and the one built from the
RTE tables (the benchmark's), each against its own fixture."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = {"synthetic": GOLDEN / "detector_scale_route.npz",
            "rte": GOLDEN / "detector_scale_route_rte.npz"}
ACTIVE = {"synthetic": 125, "rte": 38}


@pytest.fixture(scope="module", params=["synthetic", "rte"])
def event(device, request):
    from lighthit_amd.synthetic import workload_route
    pack = request.param
    want = dict(np.load(FIXTURES[pack], allow_pickle=False))
    want["pack"] = pack
    print(f"detector scale: 2304 OMs, 50000 elements, plain form, {pack} pack", flush=True)
    workload, kernel, route = workload_route("detector", device=device, pack=pack)
    assert len(workload.elements) == int(want["elements"])
    response = route(workload.elements)
    return workload, kernel, route, response, want


def test_active_set_and_levels(event):
    workload, kernel, route, response, want = event
    assert tuple(route.tables()["fold"].bands) == tuple(want["bands"])
    threshold = float(kernel.config.threshold_pe)
    total = want["charge"].sum(axis=1)
    differ = response.active != want["active"]
    # only OMs within the charge tolerance of the threshold may differ
    assert np.all(np.abs(total[differ] - threshold) <= 1e-5 * threshold), np.flatnonzero(differ)
    assert want["active"].sum() == ACTIVE[want["pack"]]
    assert np.array_equal(response.metadata["ge2_charge_level"], want["level"].tolist())
    assert np.array_equal(response.metadata["prompt_order1_level"], want["order1_level"].tolist())
    assert response.metadata["ge2_predictor"]["fallback_to_exact"] == bool(want["fallback"])
    np.testing.assert_allclose(response.time_origin_ns, want["origins"], rtol=0, atol=1e-9)


def test_charges_per_order(event):
    workload, kernel, route, response, want = event
    threshold = float(kernel.config.threshold_pe)
    got, ref = response.charge_components_pe, want["charge"]
    above = ref.sum(axis=1) >= threshold
    np.testing.assert_allclose(got[above].sum(axis=1), ref[above].sum(axis=1), rtol=1e-5)
    for order, tolerance in ((0, 1e-5), (1, 1e-6), (2, 1e-5)):
        scale = np.abs(ref[:, order]).max()
        err = np.abs(got[:, order] - ref[:, order])
        assert np.all(err <= tolerance * np.abs(ref[:, order]) + 1e-7 * scale), (
            order, float((err / np.maximum(np.abs(ref[:, order]), 1e-300)).max()))


def test_bins_and_spectrum_of_active_oms(event):
    workload, kernel, route, response, want = event
    ids = want["active_ids"]
    assert np.all(response.active[ids])
    for order, tolerance in ((0, 1e-4), (1, 1e-6), (2, 2e-5)):
        ref = want["bins"][:, :, order].astype(float)
        scale = np.maximum(np.abs(ref).sum(axis=1), 1e-12)
        err = np.abs(response.components_pe[ids, :, order] - ref).sum(axis=1) / scale
        assert err.max() < tolerance, (order, err.max())
    spectrum = want["spectrum"].astype(complex)
    err = np.abs(response.spectrum_ge2_pe[:, ids] - spectrum).max() / np.abs(spectrum).max()
    assert err < 1e-5


def test_precision_form_within_1e3_of_fp64(event, device):
    from dataclasses import replace

    from lighthit_amd.route import ShowerTransport
    workload, kernel, route, response, want = event
    print("detector scale: precision form", flush=True)
    mixed = ShowerTransport(route.kernel, route.data, config=replace(route.config,
                                                                     precision="mixed"),
                            device=device)
    got = mixed(workload.elements)
    ids = want["active_ids"]
    ref = want["charge"][ids].sum(axis=1)
    err = np.abs(got.charge_pe[ids] / ref - 1)
    print(f"precision form, detector scale: max rel charge error {err.max():.2e}")
    assert err.max() <= 1e-3
    bins = got.components_pe[ids].sum(axis=2)
    ref_bins = want["bins"].astype(float).sum(axis=2)
    assert np.max(np.abs(bins - ref_bins).sum(axis=1) / np.abs(ref_bins).sum(axis=1)) <= 1e-3
