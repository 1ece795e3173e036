"""GPU parity of the near-field clamp (lh_near_field) and of the route on an
event that passes within the tables' radial minimum of an OM."""
import warnings

import numpy as np
import pytest

from route_cases import demonstration_route, near_event

pytestmark = pytest.mark.gpu


def _close(got, want, tolerance):
    scale = np.abs(want).max()
    assert scale > 0
    err = np.abs(np.asarray(got) - np.asarray(want)).max() / scale
    assert err <= tolerance, err


def test_near_field_matches_the_oracle(device):
    from lighthit_amd.ge2 import DevicePack
    from lighthit_amd.near_field import near_field_ge2
    from oracle import near_field as oracle_near
    _, kernel, pack, route = demonstration_route(device)
    w, elements = near_event()
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, 8, np.arange(65))
    origins = np.linspace(-5.0, 40.0, len(w.detector))
    got = near_field_ge2(kernel, elements, DevicePack(bundle, device), origins, device=device)
    oms, charge, spectrum = oracle_near.near_field_ge2(kernel, elements, bundle, origins)
    assert np.array_equal(got.oms, oms) and len(oms) >= 1
    # FP64 on both sides; the coefficients are complex64 in both
    _close(got.charge_pe, charge, 1e-9)
    _close(got.spectrum_pe, spectrum, 1e-9)
    assert got.metadata["elements"] > 0


def test_route_with_a_near_pass_matches_the_oracle_route(device):
    from oracle import route as oracle_route
    _, kernel, pack, route = demonstration_route(device)
    w, elements = near_event()
    with pytest.warns(RuntimeWarning, match="near-field clamp"):
        response = route(elements)
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, kernel.config.source_degree,
                         np.arange(65))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        want = oracle_route.transport_shower(kernel, elements, bundle, None,
                                             config=route.config)
    assert len(want["near_oms"]) >= 1
    record = response.metadata["near_field"][0]
    assert record["oms"] == want["near_oms"].tolist()
    got = response.charge_components_pe
    assert np.array_equal(response.active, want["active"])
    lit = want["charge"][:, 0] > 1e-6
    np.testing.assert_allclose(got[lit, 0], want["charge"][lit, 0], rtol=1e-5)
    np.testing.assert_allclose(got[:, 1], want["charge"][:, 1], rtol=1e-6,
                               atol=1e-9 * want["charge"][:, 1].max())
    np.testing.assert_allclose(got[:, 2], want["charge"][:, 2], rtol=1e-5,
                               atol=1e-6 * np.abs(want["charge"][:, 2]).max())
    active = np.flatnonzero(response.active)
    ref = want["bins"][active, :, 2]
    err = (np.abs(response.components_pe[active, :, 2] - ref).sum(axis=1)
           / np.maximum(np.abs(ref).sum(axis=1), 1e-300))
    assert err.max() < 2e-5
    err = np.abs(response.spectrum_ge2_pe - want["spectrum"]).max() / np.abs(
        want["spectrum"]).max()
    assert err < 1e-5
