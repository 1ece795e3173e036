"""CPU pins of the order-1 oracle (grouped pencils, screen, fused bands)."""
from dataclasses import replace

import numpy as np
import pytest

from lighthit_amd.model import acceptance_alpha
from lighthit_amd.sources import SourcePose, SpectralLightElements, SyntheticShower
from lighthit_amd.synthetic import plain_small
from oracle import order1 as oracle


def _elements(start, direction, length=0.01, c0=1.0e6, beta=0.999, t0=0.0):
    start = np.atleast_2d(np.asarray(start, float))
    direction = np.atleast_2d(np.asarray(direction, float))
    n = len(start)
    length = np.full(n, length)
    return SpectralLightElements(start, direction, length, np.full(n, c0),
                                 np.full(n, -c0 / beta ** 2), np.full(n, beta),
                                 np.full(n, t0), t0 + length / (beta * 0.299792458),
                                 np.arange(n), np.arange(n), 1.35, None, "spectral")


def test_demonstration_event_total_matches_the_published_value():
    """README.md:133 quotes 50.68 pe of order 1 for the demonstration shower."""
    w = plain_small()
    d = w.detector
    tables = oracle.spectral_tables(w.medium, d, w.wavelength)
    el = w.elements
    fastest = float(np.min(tables["group_index"]))
    first = int(np.argmin(el.start_ns))
    origins = el.start_ns[first] + np.linalg.norm(d.positions_m - el.start_m[first],
                                                  axis=1) / (0.299792458 / fastest)
    q1, bins, level = oracle.spectral_order1(
        el, d.positions_m, (-d.orientations), d.effective_area_m2, origins, np.arange(len(d)), tables,
        acceptance_alpha(d), w.medium.g, w.edges_ns, w.threshold_pe, np.zeros(len(d)), None)
    assert abs(q1.sum() - 50.68) <= 0.005
    assert np.all(level == 2)
    # bins hold the light inside the 5 ns window (the front ramp is a
    # mass-normalised histogram of the same step, hence the 1e-6 allowance)
    assert 0.98 * q1.sum() < bins.sum() <= q1.sum() * (1 + 1e-6)


def test_pencils_conserve_the_ring_moments():
    shower = SyntheticShower.gaussian(SourcePose([3., 1., 2.], [0.3, 0.2, 0.93]),
                                      charged_track_length_m=30., elements=300, seed=4)
    el = shower.elements
    cone = 1.0 / (el.beta * 1.34)
    pencils = oracle.compress(el, cone, 0.5, 32, 48)
    assert np.isclose(pencils["w0"].sum(), el.coefficient0.sum(), rtol=1e-12)
    assert np.isclose(pencils["w2"].sum(), el.coefficient2.sum(), rtol=1e-12)
    # first moments: every ring node sits at its sub-segment midpoint
    a_step = 0.25
    n_a = np.maximum(1, np.ceil(el.length_m / a_step))
    mid = np.zeros(3)
    for i in range(len(el)):
        frac = (np.arange(n_a[i]) + 0.5) / n_a[i]
        points = el.start_m[i] + (frac * el.length_m[i])[:, None] * el.direction[i]
        mid += el.coefficient0[i] * points.mean(axis=0)
    got = (pencils["w0"][:, None] * pencils["position"]).sum(axis=0)
    np.testing.assert_allclose(got, mid, rtol=1e-10)
    # mean direction of each ring is its cone axis scaled by the cone cosine
    dirs = (pencils["w0"][:, None] * pencils["direction"]).sum(axis=0)
    assert np.linalg.norm(dirs) > 0


def test_screen_hierarchy_conserves_each_parent():
    shower = SyntheticShower.gaussian(SourcePose([3., 1., 2.], [0.3, 0.2, 0.93]),
                                      charged_track_length_m=30., elements=300, seed=4)
    el = shower.elements
    cone = 1.0 / (el.beta * 1.34)
    packed = oracle.compress_screen(el, cone, oracle.PromptSettings())
    parent = packed["hierarchy_parent"]
    roots = np.flatnonzero(parent == -2)
    child_sum = np.bincount(parent[parent >= 0], weights=packed["w0"][parent >= 0],
                            minlength=len(parent))
    np.testing.assert_allclose(child_sum[roots], packed["w0"][roots], rtol=1e-12)
    assert np.isclose(packed["w0"][roots].sum(), el.coefficient0.sum(), rtol=1e-12)
    # preorder: each root is followed by its children, next_node skips them
    for r, nxt in zip(roots, np.r_[roots[1:], len(parent)]):
        assert np.all(parent[r + 1:nxt] == r)
        assert packed["next_node"][r] == nxt


def _direct_single_scattering(element_mid, axis, cone, om, look, area, tables, g, acceptance):
    """Time-integrated single scattering of a point ring emitter, by quadrature:
    N mu_s int dphi/2pi int ds e^{-mu_t (s + R)} p_HG(cos chi) A(d . look) area / R^2."""
    axis = axis / np.linalg.norm(axis)
    helper = np.array([0., 0., 1.]) if abs(axis[2]) < 0.9 else np.array([1., 0., 0.])
    e1 = np.cross(axis, helper)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(axis, e1)
    sine = np.sqrt(1 - cone ** 2)
    phi = (np.arange(720) + 0.5) * 2 * np.pi / 720
    u = cone * axis + sine * (np.cos(phi)[:, None] * e1 + np.sin(phi)[:, None] * e2)
    sigma = 0.05
    v = np.linspace(0.0, np.arcsinh(3000.0 / sigma), 24001)
    s = sigma * np.sinh(v)
    ds = sigma * np.cosh(v) * (v[1] - v[0])
    ds[0] *= 0.5
    ds[-1] *= 0.5
    x = element_mid + s[None, :, None] * u[:, None, :]
    dvec = om - x
    R = np.linalg.norm(dvec, axis=2)
    dhat = dvec / R[..., None]
    cos_chi = np.einsum("pk,psk->ps", u, dhat)
    accept = acceptance(dhat @ look)
    total = 0.0
    for lam in range(len(tables["mus"])):
        mut = tables["mut"][lam]
        hg = (1 - g * g) / (4 * np.pi * (1 + g * g - 2 * g * cos_chi) ** 1.5)
        integrand = np.exp(-mut * (s[None, :] + R)) * hg * accept / R ** 2
        photons = tables["s0w"][lam] * 1.0e6 + tables["s2w"][lam] * (-1.0e6 / 0.999 ** 2)
        total += photons * tables["mus"][lam] * (integrand @ ds).mean()
    return area * total


@pytest.mark.parametrize("om,look", [
    ([20.0, 0.0, 3.0], [0.0, 0.0, -1.0]),
    ([0.0, -12.0, -9.0], [0.6, 0.0, 0.8]),
    ([5.0, 7.0, 25.0], [0.0, -1.0, 0.0]),
])
def test_grouped_charge_matches_a_direct_single_scattering_integral(om, look):
    w = plain_small()
    tables = oracle.spectral_tables(w.medium, w.detector, w.wavelength)
    one = {k: v[1:2] for k, v in tables.items()}
    phase = float(one["phase_index"][0])
    el = _elements([0.0, 0.0, 0.0], [0.0, 0.0, 1.0])
    cone = 1.0 / (el.beta * phase)
    alpha = acceptance_alpha(w.detector)
    pencils = oracle.compress(el, cone, 0.5, 96, 192)
    om = np.asarray(om, float)
    look = np.asarray(look, float) / np.linalg.norm(look)
    edges = np.arange(-100.0, 4000.0, 5.0)
    charge, bins, _, _ = oracle.grouped(om[None], look[None], np.array([0.1257]),
                                        np.array([0.0]), [0], pencils, one, alpha, w.medium.g,
                                        edges, oracle.PromptSettings(),
                                        (2e-3, 0.08, 0.5, 0.02))
    acceptance = w.detector.angular_acceptance
    direct = _direct_single_scattering(np.array([0, 0, 0.005]), np.array([0., 0., 1.]),
                                       float(cone[0]), om, look, 0.1257, one, w.medium.g,
                                       acceptance)
    assert charge[0] == pytest.approx(direct, rel=5e-3)
    # a planted error: dropping the HG phase normalisation would fail
    assert abs(charge[0] * 4 * np.pi - direct) > 0.5 * direct
    # the window holds all the light: bins sum to the charge
    assert bins.sum() == pytest.approx(charge[0], rel=1e-6)


def test_rigid_time_shift_leaves_bins_unchanged():
    w = plain_small()
    tables = oracle.spectral_tables(w.medium, w.detector, w.wavelength)
    alpha = acceptance_alpha(w.detector)
    el = _elements([[0., 0., 0.], [0.1, 0.0, 0.5]], [[0., 0., 1.], [0.1, 0., 1.]])
    shifted = replace(el, start_ns=el.start_ns + 1234.5, end_ns=el.end_ns + 1234.5)
    om = np.array([[15.0, 2.0, 4.0], [-9.0, 3.0, 12.0]])
    looks = np.array([[0., 0., -1.], [1., 0., 0.]])
    origins = np.array([40.0, 55.0])
    out = []
    for elements, shift in ((el, 0.0), (shifted, 1234.5)):
        cone = 1.0 / (elements.beta * 1.34)
        pencils = oracle.compress(elements, cone, 0.5, 96, 192)
        out.append(oracle.grouped(om, looks, np.full(2, 0.1257), origins + shift, [0, 1],
                                  pencils, tables, alpha, w.medium.g, w.edges_ns,
                                  oracle.PromptSettings(), (2e-3, 0.08, 0.5, 0.02)))
    np.testing.assert_allclose(out[0][0], out[1][0], rtol=1e-9)
    np.testing.assert_allclose(out[0][1], out[1][1], rtol=1e-7, atol=1e-12 * out[0][1].max())
