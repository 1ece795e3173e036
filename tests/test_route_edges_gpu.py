"""Edge cases of the whole route on the demonstration array: an event far from
every OM, a one-element event, a source below the Cherenkov threshold, and
sub-threshold elements mixed into an event."""
from dataclasses import replace

import numpy as np
import pytest

from route_cases import demonstration_route

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(device):
    w, kernel, pack, route = demonstration_route(device)
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, kernel.config.source_degree,
                         np.arange(65))
    return w, kernel, route, bundle


def _oracle(setup, elements):
    from oracle import route as oracle_route
    w, kernel, route, bundle = setup
    return oracle_route.transport_shower(kernel, elements, bundle, None,
                                         config=route.config)


def test_event_far_from_every_om(setup):
    """5 km away: nothing is active, no bins, charges match the oracle's."""
    w, kernel, route, bundle = setup
    far = replace(w.elements, start_m=w.elements.start_m + np.array([5000.0, 0.0, 0.0]))
    r = route(far)
    want = _oracle(setup, far)
    assert not r.active.any() and not want["active"].any()
    assert not r.components_pe.any()
    assert np.isfinite(r.charge_components_pe).all()
    scale = max(float(np.abs(want["charge"]).max()), 1e-300)
    assert np.abs(r.charge_components_pe - want["charge"]).max() <= 1e-5 * scale + 1e-12


def test_one_element_event_matches_the_oracle(setup):
    w, kernel, route, bundle = setup
    brightest = int(np.argmax(w.elements.coefficient0))
    one = w.elements.subset(np.arange(len(w.elements)) == brightest)
    r = route(one)
    want = _oracle(setup, one)
    assert np.array_equal(r.active, want["active"])
    for order, tolerance in ((0, 1e-5), (1, 1e-6), (2, 1e-5)):
        ref = want["charge"][:, order]
        lit = np.abs(ref) > 1e-6 * max(np.abs(ref).max(), 1e-300)
        if lit.any():
            rel = np.abs(r.charge_components_pe[lit, order] - ref[lit]) / np.abs(ref[lit])
            assert rel.max() <= tolerance, (order, rel.max())


def test_source_below_threshold_raises(setup):
    w, kernel, route, bundle = setup
    slow = replace(w.elements, beta=np.full(len(w.elements), 0.5))
    with pytest.raises(ValueError, match="below Cherenkov threshold"):
        route(slow)


def test_sub_threshold_elements_do_not_change_the_event(setup):
    """Elements that never emit are dropped: the event with them equals the
    event without them."""
    w, kernel, route, bundle = setup
    n = len(w.elements)
    extra = w.elements.subset(np.arange(n) < 50)
    extra = replace(extra, beta=np.full(len(extra), 0.5))
    from dataclasses import fields as dataclass_fields
    joined = {}
    for f in dataclass_fields(w.elements):
        value = getattr(w.elements, f.name)
        if isinstance(value, np.ndarray) and len(value) == n:
            joined[f.name] = np.concatenate([value, np.asarray(getattr(extra, f.name))])
    mixed = replace(w.elements, **joined)
    a, b = route(w.elements), route(mixed)
    assert np.array_equal(a.active, b.active)
    np.testing.assert_allclose(b.charge_components_pe, a.charge_components_pe, rtol=1e-9,
                               atol=1e-12)
