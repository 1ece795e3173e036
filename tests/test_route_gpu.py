"""GPU parity of the whole shower route against the oracle route."""
import numpy as np
import pytest

from route_cases import demonstration_route

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def demonstration(device):
    from oracle import route as oracle_route
    w, kernel, pack, route = demonstration_route(device)
    response = route(w.elements)
    tables = route.tables()
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, kernel.config.source_degree,
                         np.arange(65))
    expected = oracle_route.transport_shower(kernel, w.elements, bundle, None,
                                             config=route.config)
    return w, route, response, expected


def test_route_charges_match_the_oracle(demonstration):
    w, route, response, want = demonstration
    got = response.charge_components_pe
    assert np.array_equal(response.active, want["active"])
    lit = want["charge"][:, 0] > 1e-6
    np.testing.assert_allclose(got[lit, 0], want["charge"][lit, 0], rtol=1e-5)
    np.testing.assert_allclose(got[:, 1], want["charge"][:, 1], rtol=1e-6,
                               atol=1e-9 * want["charge"][:, 1].max())
    np.testing.assert_allclose(got[:, 2], want["charge"][:, 2], rtol=1e-5,
                               atol=1e-6 * np.abs(want["charge"][:, 2]).max())
    # the published demonstration totals of orders 0 and 1 (README.md:133)
    assert abs(got[:, 1].sum() - 50.68) < 0.01
    assert np.array_equal(response.metadata["ge2_charge_level"], want["level"].tolist())


def test_route_bins_and_spectrum_match_the_oracle(demonstration):
    w, route, response, want = demonstration
    active = np.flatnonzero(response.active)
    charge = want["charge"].sum(axis=1)[active]
    for order, tolerance in ((0, 1e-4), (1, 1e-6), (2, 2e-5)):
        got = response.components_pe[active, :, order]
        ref = want["bins"][active, :, order]
        scale = np.maximum(np.abs(ref).sum(axis=1), 1e-12 * charge)
        err = np.abs(got - ref).sum(axis=1) / scale
        assert err.max() < tolerance, (order, err.max())
    spectrum = want["spectrum"]
    err = np.abs(response.spectrum_ge2_pe - spectrum).max() / np.abs(spectrum).max()
    assert err < 1e-5
    np.testing.assert_allclose(response.time_origin_ns, want["origins"], rtol=0, atol=1e-9)


def test_response_interface(demonstration, tmp_path):
    w, route, response, want = demonstration
    n = len(w.detector)
    assert response.charge_components_pe.shape == (n, 3)
    assert response.components_pe.shape == (n, len(w.edges_ns) - 1, 3)
    assert response.spectrum_ge2_pe.shape == (65, n)
    assert response.metadata["route"] == "lighthit.shower"
    np.testing.assert_allclose(response.charge_pe, response.charge_components_pe.sum(axis=1))
    prompt = response.select("0+1")
    np.testing.assert_allclose(prompt.bins_pe, response.components_pe[:, :, :2].sum(axis=2))
    path = response.save(tmp_path / "event.npz")
    with np.load(path, allow_pickle=False) as saved:
        np.testing.assert_array_equal(saved["active"], response.active)
        assert "lighthit/hybrid-response/1" in str(saved["metadata"])


def test_repeated_calls_are_identical(demonstration):
    w, route, response, want = demonstration
    again = route(w.elements)
    np.testing.assert_array_equal(again.charge_components_pe, response.charge_components_pe)
    np.testing.assert_array_equal(again.components_pe, response.components_pe)


def test_receiver_subset_computes_every_listed_om(demonstration):
    w, route, response, want = demonstration
    ids = np.array([3, 7, 11, 40])
    part = route(w.elements, receivers=ids)
    assert part.charge_components_pe.shape == (4, 3)
    assert np.all(part.active)
    np.testing.assert_allclose(part.charge_components_pe[:, :2],
                               response.charge_components_pe[ids, :2], rtol=1e-6)


def _oracle_window(geometry_kernel, item, low, high, efficiency):
    from lighthit_amd.route import spectral_attenuation
    from oracle import route as oracle_route
    d = geometry_kernel.detector
    radius, log_factor = spectral_attenuation(geometry_kernel)
    peak = float(np.max(np.abs(np.asarray(d.angular_acceptance(np.linspace(-1, 1, 257)),
                                          float))))
    shape = tuple(item.channels.shape)
    ch = item.channels.cpu().numpy().reshape(shape[0], int(np.prod(shape[1:])))[:, 0]
    photons = np.sqrt(4 * np.pi) * np.abs(ch).astype(float) / efficiency
    return oracle_route.window_report(item.to_points_m(), photons, d.positions_m,
                                      d.effective_area_m2, low, high, radius, log_factor,
                                      peak, efficiency)


def test_window_report_matches_the_oracle(demonstration):
    from lighthit_amd.route import window_report
    w, route, response, want = demonstration
    kernel = route.kernel
    geometry = route.window_geometry(kernel)
    low, high = kernel.config.radial_range_m
    efficiency = np.asarray(kernel.detector.spectral_weight(kernel.wavelength.wavelength_nm))
    sources = route._compiled["sources"][0]
    bands = route.tables()["fold"].bands
    for band, item in zip(bands, sources):
        got = [t.cpu().numpy() for t in window_report(geometry, item, low, high,
                                                      float(efficiency[band]))]
        ref = _oracle_window(kernel, item, low, high, float(efficiency[band]))
        for a, b in zip(got[:3], ref[:3]):
            assert np.array_equal(a.astype(np.int64), b)
        # the channel magnitude is a float32 hypot on both sides (one ulp apart)
        np.testing.assert_allclose(got[3], ref[3], rtol=1e-6, atol=1e-300)
        assert np.all(got[0] + got[1] + got[2] == item.points.shape[0])


def test_window_report_every_branch(device):
    """Points below, inside and beyond the window of every OM, with an empty band."""
    import torch
    from lighthit_amd.axial import AxisFrame
    from lighthit_amd.compile import CompiledBand
    from lighthit_amd.route import RouteConfig, RouteKernel, window_report, WindowGeometry
    from lighthit_amd.synthetic import plain_small
    w = plain_small()
    kernel = RouteKernel(w.medium, w.detector, RouteConfig())
    geometry = WindowGeometry(kernel, device)
    rng = np.random.default_rng(5)
    centre = w.detector.positions_m.mean(axis=0)
    axis = np.array([0.2, -0.3, 0.93]) / np.linalg.norm([0.2, -0.3, 0.93])
    first = np.cross(axis, [1.0, 0, 0])
    first /= np.linalg.norm(first)
    frame = AxisFrame(centre, axis, first, np.cross(axis, first))
    for count in (0, 1, 700):
        pts = np.column_stack([rng.normal(scale=40, size=count), rng.normal(scale=40, size=count),
                               rng.uniform(-400, 400, size=count)])
        ch = (rng.normal(size=(count, 9, 3)) + 1j * rng.normal(size=(count, 9, 3)))
        item = CompiledBand(frame, torch.as_tensor(pts, device=device),
                            torch.as_tensor(ch.astype(np.complex64), device=device), 2, 2, 0.0)
        got = [t.cpu().numpy() for t in window_report(geometry, item, 2.0, 150.0, 0.7)]
        ref = _oracle_window(kernel, item, 2.0, 150.0, 0.7)
        for a, b in zip(got[:3], ref[:3]):
            assert np.array_equal(a.astype(np.int64), b)
        np.testing.assert_allclose(got[3], ref[3], rtol=1e-6, atol=1e-300)
        if count == 700:
            assert ref[1].min() > 0 and ref[0].min() > 0


def test_served_results_match_alone(demonstration, device):
    """Concurrent calls on several streams give every event the result it has alone."""
    from lighthit_amd.route import ResidentShower
    from dataclasses import replace

    from lighthit_amd.serve import ShowerServer
    w, route, response, want = demonstration
    # three different events: the demonstration shower and two shifted copies
    events = [w.elements]
    for shift in ([4.0, -3.0, 2.0], [-6.0, 5.0, -1.0]):
        events.append(replace(w.elements, start_m=w.elements.start_m + np.asarray(shift)))
    alone = [route(ResidentShower(e, device)) for e in events]
    with ShowerServer(route, workers=3, device=device) as server:
        futures = [server.submit(ResidentShower(events[i % 3], device)) for i in range(9)]
        mixed = [f.result() for f in futures]
    for i, r in enumerate(mixed):
        a = alone[i % 3]
        assert np.array_equal(r.charge_components_pe, a.charge_components_pe)
        assert np.array_equal(r.components_pe, a.components_pe)
        assert np.array_equal(r.active, a.active)
    assert all(f.latency_s > 0 for f in futures)


def test_process_served_results_match_alone(demonstration, device):
    """Events served by two worker processes, each with its own route on the
    same GPU, carry the result the parent's route gives them alone."""
    import functools
    from dataclasses import replace

    from route_cases import demonstration_route_only

    from lighthit_amd.serve import ProcessShowerServer
    w, route, response, want = demonstration
    events = [w.elements]
    for shift in ([4.0, -3.0, 2.0], [-6.0, 5.0, -1.0]):
        events.append(replace(w.elements, start_m=w.elements.start_m + np.asarray(shift)))
    alone = [route(e) for e in events]
    with ProcessShowerServer(functools.partial(demonstration_route_only, "cuda:0"), processes=2,
                             detector=route.kernel.detector, device="cuda:0",
                             threads_per_process=2) as server:
        # host arrays for the first three, sources staged in device memory after
        staged = [server.stage(e) for e in events]
        futures = [server.submit(events[i % 3] if i < 3 else staged[i % 3]) for i in range(6)]
        mixed = [f.result(timeout=300) for f in futures]
        for handle in staged:
            server.release(handle)
    for i, r in enumerate(mixed):
        a = alone[i % 3]
        assert np.array_equal(r.charge_components_pe, a.charge_components_pe)
        assert np.array_equal(r.components_pe, a.components_pe)
        assert np.array_equal(r.spectrum_ge2_pe, a.spectrum_ge2_pe)
        assert np.array_equal(r.active, a.active)
        assert np.array_equal(r.time_origin_ns, a.time_origin_ns)
    assert all(f.latency_s > 0 for f in futures)
