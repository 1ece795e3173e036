"""The GPU route with the order >=2 pack built from the RTE tables: the published
demonstration result (README.md:133: 34 active OMs, 22.5 pe of order >=2), and
parity with the oracle route on the same pack, in both precision forms."""
from dataclasses import replace

import numpy as np
import pytest

from route_cases import demonstration_route

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def real(device):
    from oracle import route as oracle_route

    from lighthit_amd.route import ShowerTransport
    w, kernel, pack, route = demonstration_route(device, tables=True)
    plain = route(w.elements)
    mixed = ShowerTransport(kernel, pack, device=device,
                            config=replace(route.config, precision="mixed"))(w.elements)
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, kernel.config.source_degree,
                         np.arange(65))
    want = oracle_route.transport_shower(kernel, w.elements, bundle, None, config=route.config)
    return w, plain, mixed, want


def test_published_order2_total_and_active_set(real):
    w, plain, mixed, want = real
    totals = plain.charge_components_pe.sum(axis=0)
    print("GPU route with RTE tables: orders", totals, "active", int(plain.active.sum()))
    assert int(plain.active.sum()) == 34
    assert abs(totals[2] - 22.5) < 0.05
    assert abs(totals[1] - 50.68) < 0.01


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_route_with_tables_matches_the_oracle(real, form):
    w, plain, mixed, want = real
    r = plain if form == "plain" else mixed
    assert np.array_equal(r.active, want["active"])
    # the bar of the plain form: charge_pe within 1e-5 on every OM above threshold
    total = want["charge"].sum(axis=1)
    above = r.active
    rel = np.abs(r.charge_pe[above] - total[above]) / total[above]
    assert rel.max() <= 1e-5, (form, rel.max())
    # per order, as the demonstration route's parity test
    for order, rtol in ((0, 1e-5), (1, 1e-6), (2, 1e-5)):
        ref = want["charge"][:, order]
        np.testing.assert_allclose(r.charge_components_pe[:, order], ref, rtol=rtol,
                                   atol=1e-6 * np.abs(ref).max())
    active = np.flatnonzero(r.active)
    charge = want["charge"].sum(axis=1)[active]
    bars = ((0, 1e-4), (1, 1e-6), (2, 2e-5 if form == "plain" else 1e-3))
    for order, tolerance in bars:
        got = r.components_pe[active, :, order]
        ref = want["bins"][active, :, order]
        scale = np.maximum(np.abs(ref).sum(axis=1), 1e-12 * charge)
        err = np.abs(got - ref).sum(axis=1) / scale
        assert err.max() < tolerance, (form, order, err.max())
