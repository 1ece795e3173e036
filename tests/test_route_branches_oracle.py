"""CPU pins of the oracle route's delayed-light, local-origin and fallback
branches (shower/api.py:356-375, 405-416, 468-477), from invariants the
reference states for them."""
import numpy as np
import pytest

from route_cases import delayed_pair, long_source


@pytest.fixture(scope="module")
def setup():
    from lighthit_amd.route import RouteConfig, RouteKernel, ShowerConfig
    from lighthit_amd.synthetic import plain_small, synthetic_pack
    w = plain_small()
    kernel = RouteKernel(w.medium, w.detector, RouteConfig(relative_time_edges_ns=w.edges_ns))
    bundle = synthetic_pack(w.medium, kernel.wavelength)[2]
    return w, kernel, bundle, ShowerConfig(delayed_cell_ns=100.0)


def _route(setup, elements, **options):
    from oracle import route as oracle_route
    w, kernel, bundle, config = setup
    return oracle_route.transport_shower(kernel, elements, bundle, config=config, **options)


@pytest.fixture(scope="module")
def alone_and_pair(setup):
    first, both = delayed_pair()
    alone = _route(setup, first)
    pair = _route(setup, both, time_origin_ns=alone["origins"])
    return alone, pair


def test_delayed_copy_is_the_same_light_on_a_shifted_clock(setup, alone_and_pair):
    """tests/test_shower_route.py:200-228: a copy emitted 250 ns (50 bins) later is one
    delayed group; it doubles the order >=2 charge and adds the original waveform read
    on edges seen 250 ns later, which away from the window start is the original
    shifted by 50 bins."""
    from oracle import readout as oracle_readout
    w, kernel, bundle, config = setup
    alone, pair = alone_and_pair
    assert pair["groups_in_window"] == 1 and not alone["local_origin"]
    act = np.flatnonzero(alone["active"] & (alone["level"] == 2))
    assert len(act) > 10 and np.all(pair["active"][act])
    np.testing.assert_allclose(pair["charge"][act, 2], 2 * alone["charge"][act, 2], rtol=1e-6)
    edges = kernel.config.relative_time_edges_ns
    copy = oracle_readout.route_readout(kernel.config.omega_per_ns, alone["spectrum"][:, act],
                                        alone["origins"][act],
                                        kernel.config.readout_omega_per_ns, edges - 250.0)
    scale = np.abs(pair["bins"][act, :, 2]).max()
    np.testing.assert_allclose(copy[:, 50:], alone["bins"][act, :-50, 2], rtol=0,
                               atol=1e-9 * scale)
    expected = alone["bins"][act, :, 2] + copy
    # (the copy is compiled on its own time lattice: 0.02 ns ticks from its first step)
    assert np.abs(pair["bins"][act, :, 2] - expected).max() <= 2e-6 * scale
    # direct light of the copy arrives 250 ns later on the common clock (order 1
    # merges the two copies' ring nodes into shared pencils, so it is no sum of two)
    late = pair["bins"][act, :, 0] - alone["bins"][act, :, 0]
    np.testing.assert_allclose(late[:, 50:], alone["bins"][act, :-50, 0], rtol=0,
                               atol=1e-6 * alone["bins"][act, :, 0].max())


def test_long_source_takes_local_origins(setup):
    """A source longer than the radial range reads every OM on its own clock: the
    earliest arrival over the segments at the OM centre, so no direct light lands
    more than the glass radius's flight time (0.2 m x n_g / c < 1 ns) before it."""
    from oracle import host
    w, kernel, bundle, config = setup
    source = long_source()
    out = _route(setup, source)
    assert out["local_origin"]
    group = float(np.min(kernel.medium.sample(kernel.wavelength.wavelength_nm)["group_index"]))
    np.testing.assert_allclose(out["origins"], host.earliest_arrival(
        w.detector.positions_m, source, group), rtol=0, atol=1e-9)
    first = int(np.argmin(source.start_ns))
    global_origin = source.start_ns[first] + np.linalg.norm(
        w.detector.positions_m - source.start_m[first], axis=1) * group / 0.299792458
    assert np.all(out["origins"] <= global_origin + 1e-9)
    assert np.any(out["origins"] < global_origin - 50.0)
    edges = kernel.config.relative_time_edges_ns
    early = edges[1:] <= -1.0
    active = np.flatnonzero(out["active"])
    assert len(active) and not out["bins"][active][:, early, 0].any()


def test_forced_fallback_computes_every_reachable_om_exactly(setup, alone_and_pair):
    """tests/test_shower_route.py:161-178: a prediction 1e6 too small fails its own
    margin, and every reachable OM gets the exact zero-frequency charge."""
    first, _ = delayed_pair()
    alone, _ = alone_and_pair
    forced = _route(setup, first, predictor_scale=1e-6)
    assert forced["fallback"] and not alone["fallback"]
    levels = forced["level"]
    assert np.all(levels[levels > 0] == 2)
    both = (alone["level"] == 2) & (levels == 2)
    np.testing.assert_array_equal(forced["charge"][both, 2], alone["charge"][both, 2])
