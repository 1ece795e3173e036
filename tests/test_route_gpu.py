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
    expected = oracle_route.transport_shower(kernel, w.elements, bundle, tables["fold"],
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
