"""GPU parity of the route's delayed-light, local-origin and fallback branches,
and of lh_segment_bounds, against the oracle (oracle/route.py, oracle/host.py)."""
import numpy as np
import pytest

from route_cases import delayed_pair, long_source

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(device):
    from lighthit_amd.fold import KernelPack
    from lighthit_amd.route import RouteConfig, RouteKernel, ShowerConfig, ShowerTransport
    from lighthit_amd.synthetic import plain_small, synthetic_pack
    w = plain_small()
    kernel = RouteKernel(w.medium, w.detector, RouteConfig(relative_time_edges_ns=w.edges_ns))
    pack = KernelPack()
    pack.add(*synthetic_pack(w.medium, kernel.wavelength))
    config = ShowerConfig(delayed_cell_ns=100.0)
    route = ShowerTransport(kernel, pack, config=config, device=device)
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, kernel.config.source_degree,
                         np.arange(65))
    return w, kernel, route, bundle, config


def _oracle(setup, elements, **options):
    from oracle import route as oracle_route
    w, kernel, route, bundle, config = setup
    return oracle_route.transport_shower(kernel, elements, bundle, config=config, **options)


def _compare(response, want):
    """The route's own bars: identical active set and levels; charges per order;
    bins per order against the OM's bin L1."""
    assert np.array_equal(response.active, want["active"])
    assert np.array_equal(response.metadata["ge2_charge_level"], want["level"].tolist())
    got = response.charge_components_pe
    for order, tolerance in ((0, 1e-5), (1, 1e-6), (2, 1e-5)):
        ref = want["charge"][:, order]
        assert np.all(np.abs(got[:, order] - ref)
                      <= tolerance * np.abs(ref) + 1e-6 * np.abs(ref).max()), order
    active = np.flatnonzero(response.active)
    for order, tolerance in ((0, 1e-4), (1, 1e-6), (2, 2e-5)):
        ref = want["bins"][active, :, order]
        scale = np.maximum(np.abs(ref).sum(axis=1), 1e-12)
        err = np.abs(response.components_pe[active, :, order] - ref).sum(axis=1) / scale
        assert err.max() < tolerance, (order, err.max())
    np.testing.assert_allclose(response.time_origin_ns, want["origins"], rtol=0, atol=1e-9)


def test_delayed_copy_matches_the_oracle(setup):
    first, both = delayed_pair()
    w, kernel, route, bundle, config = setup
    print("branch: delayed copy, 2 x 300 elements", flush=True)
    alone = route(first)
    pair = route(both, time_origin_ns=alone.time_origin_ns)
    assert pair.metadata["delayed_light"]["groups_in_window"] == 1
    want = _oracle(setup, both, time_origin_ns=alone.time_origin_ns)
    assert want["groups_in_window"] == 1
    _compare(pair, want)
    # the reference's invariant on the GPU: the copy doubles the exact order >=2
    act = np.flatnonzero(alone.active & (np.asarray(alone.metadata["ge2_charge_level"]) == 2))
    np.testing.assert_allclose(pair.charge_components_pe[act, 2],
                               2 * alone.charge_components_pe[act, 2], rtol=1e-6)


def test_long_source_local_origins_match_the_oracle(setup):
    source = long_source()
    w, kernel, route, bundle, config = setup
    print("branch: long source, 2 x 300 elements 200 m apart", flush=True)
    response = route(source)
    want = _oracle(setup, source)
    assert want["local_origin"]
    assert response.metadata["time_origin"] == "earliest local segment and fastest group flight"
    _compare(response, want)


def test_forced_fallback_matches_the_oracle(setup, monkeypatch):
    """tests/test_shower_route.py:161-178: the collapsed predictor 1e6 too small."""
    import lighthit_amd.route as product_route
    from dataclasses import replace
    first, _ = delayed_pair()
    w, kernel, route, bundle, config = setup
    original = product_route.collapse

    def tiny(sources, cell_m):
        return [replace(item, channels=item.channels * 1e-6)
                for item in original(sources, cell_m)]
    monkeypatch.setattr(product_route, "collapse", tiny)
    print("branch: forced predictor fallback, 300 elements", flush=True)
    response = route(first)
    predictor = response.metadata["ge2_predictor"]
    assert predictor["fallback_to_exact"] is True
    levels = np.asarray(response.metadata["ge2_charge_level"])
    assert np.all(levels[levels > 0] == 2)
    want = _oracle(setup, first, predictor_scale=1e-6)
    assert want["fallback"]
    _compare(response, want)


@pytest.mark.parametrize("count", [1, 7, 13, 53])
def test_segment_bounds_match_fp64(device, count):
    """lh_segment_bounds against min_distance_to_segments and earliest_element_arrival
    (_prompt_numba.py:2221-2243, _shower_basis.py:51-78), OM counts that are not a
    multiple of the kernel's 8 OMs per workgroup."""
    from lighthit_amd.route import segment_bounds
    from oracle import host
    source = long_source(elements=500)
    rng = np.random.default_rng(count)
    positions = rng.uniform(-250, 250, (count, 3)) + source.start_m.mean(axis=0)
    distance, earliest = segment_bounds(positions, source, 1.367 / 0.299792458, device)
    np.testing.assert_allclose(distance, host.min_distance(positions, source), rtol=1e-12)
    np.testing.assert_allclose(earliest, host.earliest_arrival(positions, source, 1.367),
                               rtol=1e-12)
