"""CPU pins of the stage (a) oracle: first-entry direct light."""
from dataclasses import replace

import numpy as np
import pytest

from lighthit_amd.model import DetectorArray, SpectralMedium, SphereSurface, acceptance_alpha
from lighthit_amd.optics import prepare_optics
from lighthit_amd.sources import SpectralLightElements
from oracle import order0

C = 0.299792458


def _element(length, beta=1.0, coefficient=1e8):
    return SpectralLightElements(np.zeros((1, 3)), np.array([[0., 0., 1.]]), np.array([length]),
                                 np.array([coefficient * length]),
                                 np.array([-coefficient * length / beta ** 2]), np.array([beta]),
                                 np.zeros(1), np.array([length / (beta * C)]), np.array([0]),
                                 np.array([0]), cone_model='spectral')


def _run(detector, medium, elements, alpha, **kw):
    optics = prepare_optics(medium, detector)
    radii = detector.receiver_surface.radii(len(detector))
    return order0.first_entry(detector.positions_m, (-detector.orientations), detector.effective_area_m2,
                              radii, elements, optics, alpha, **kw)


def test_axial_constant_phase_closed_form():
    # tests/test_ballistic_first_entry.py:27-47 of the reference: OM on the
    # element axis, constant n: every azimuth hits, the spectrum integrates in
    # closed form and the azimuth integral is trivial.
    medium = SpectralMedium([350., 610.], [.7, .7], [0., 0.], [1.35, 1.35], [1.4, 1.4], g=.9)
    detector = DetectorArray([[0., 0., .25]], [[0., 0., -1.]], .05, lambda x: np.ones_like(x),
                             lambda w: np.full_like(w, .3), receiver_surface=SphereSurface(.2))
    length = .02
    elements = _element(length)
    alpha = acceptance_alpha(detector, acceptance_degree=0)
    x, w = np.polynomial.legendre.leggauss(96)
    a = (1 + x) * length / 2
    along = .25 - a
    cosine = 1 / 1.35
    entry = cosine * along - np.sqrt(.2 ** 2 - (1 - cosine ** 2) * along ** 2)
    factor = 1e8 * (1 - cosine ** 2) * .3 * (1 / 350 - 1 / 610) * .05 / (np.pi * .2 ** 2)
    expected = factor * np.dot(w, np.exp(-.7 * entry)) * length / 2
    endpoints = np.array([0., length])
    r = .25 - endpoints
    path = cosine * r - np.sqrt(.2 ** 2 - (1 - cosine ** 2) * r ** 2)
    time = endpoints / C + 1.4 * path / C
    edges = np.array([-5., time.min() - 1e-8, time.max() + 1e-8, 5.])
    charge, bins = _run(detector, medium, elements, alpha, origins=[0.], edges=edges)
    np.testing.assert_allclose(charge, [expected], rtol=2e-8)
    np.testing.assert_allclose(bins[0], [0., expected, 0.], rtol=2e-8, atol=1e-11)


def test_axial_cone_that_misses_has_zero_charge():
    medium = SpectralMedium([350., 610.], [.7, .7], [0., 0.], [1.35, 1.35], [1.4, 1.4], g=.9)
    detector = DetectorArray([[0., 0., 2.]], [[0., 0., -1.]], .05, lambda x: np.ones_like(x),
                             lambda w: np.full_like(w, .3), receiver_surface=SphereSurface(.2))
    charge, _ = _run(detector, medium, _element(.02), acceptance_alpha(detector, acceptance_degree=0))
    np.testing.assert_array_equal(charge, 0.)


def test_emission_crossing_the_glass_is_rejected():
    medium = SpectralMedium([350., 610.], [.7, .7], [0., 0.], [1.35, 1.35], [1.4, 1.4], g=.9)
    detector = DetectorArray([[.1, 0., .01]], [[0., 0., -1.]], .05, lambda x: np.ones_like(x),
                             lambda w: np.full_like(w, .3), receiver_surface=SphereSurface(.2))
    with pytest.raises(ValueError, match='intersects the glass'):
        _run(detector, medium, _element(.02), acceptance_alpha(detector, acceptance_degree=0))


def _independent_rays(medium, centre, look, radius, accept, beta, length, n=64):
    """Brute-force (wavelength, source point, azimuth) integral of the first-entry charge.

    Written from the physics: every wavelength node, every emission point on
    the element and every azimuth on the Cherenkov cone that reaches the
    sphere, weighted by acceptance and extinction over the entry distance.
    """
    gx, gw = np.polynomial.legendre.leggauss(n)
    lam = 480 + 130 * gx
    sampled = medium.sample(lam)
    cosine = 1 / (beta * sampled['phase_index'])
    sine = np.sqrt((1 - cosine) * (1 + cosine))
    phi_node = np.sin(np.pi * gx / 2)
    phi_weight = np.pi / 2 * np.cos(np.pi * gx / 2) * gw
    total = 0.
    for a, weight_a in zip((1 + gx) * length / 2, gw * length / 2):
        w = centre - np.array([0., 0., a])
        outside = w @ w - radius ** 2
        half = np.arccos(np.clip((np.sqrt(outside) - cosine * w[2]) / (sine * w[0]), -1, 1))
        phi = half[:, None] * phi_node
        dx, dy = sine[:, None] * np.cos(phi), sine[:, None] * np.sin(phi)
        dot = dx * w[0] + cosine[:, None] * w[2]
        entry = outside / (dot + np.sqrt(np.maximum(0., dot ** 2 - outside)))
        angular = ((accept(dx * look[0] + dy * look[1] + cosine[:, None] * look[2])
                    + accept(dx * look[0] - dy * look[1] + cosine[:, None] * look[2]))
                   * np.exp(-sampled['absorption_per_m'][:, None] * entry))
        angular = (angular * phi_weight).sum(1) * half / (2 * np.pi)
        total += weight_a * np.dot(130 * gw, 1e8 * sine ** 2 / lam ** 2 * .3 * angular)
    return total


@pytest.mark.parametrize('offset_m', [0., -.12])
def test_dispersive_tilted_receiver_matches_independent_rays(offset_m):
    # tests/test_ballistic_first_entry.py:77-123 of the reference.
    beta, length, radius = .9, .05, .126
    medium = SpectralMedium([350., 610.], [.07, .04], [0., 0.], [1.38, 1.32], [1.43, 1.36], g=.9)
    look = np.array([.6, .3, np.sqrt(.55)])
    centre = np.array([2., 0., 2 / np.sqrt((beta * 1.35) ** 2 - 1) + length / 2 + offset_m])

    def accept(x):
        return 1 + .3 * x + .1 * x * x + .02 * x * x * x
    detector = DetectorArray([centre], [-look], .05, accept, lambda w: np.full_like(w, .3),
                             receiver_surface=SphereSurface(radius))
    elements = replace(_element(length, beta), coefficient0=np.array([1e8 * length]),
                       coefficient2=np.array([-1e8 * length / beta ** 2]))
    alpha = acceptance_alpha(detector, acceptance_degree=3)
    # The reference's independent integral integrates both azimuth signs over
    # the half range, so it counts the cone twice: divide by two.
    expected = 0.5 * _independent_rays(medium, centre, look, radius, accept, beta, length) \
        * .05 / (np.pi * radius ** 2)
    charge, bins = _run(detector, medium, elements, alpha, origins=[0.],
                        edges=np.arange(-5., 40., .25))
    np.testing.assert_allclose(charge, [expected], rtol=2e-6, atol=1e-10)
    np.testing.assert_allclose(bins.sum(), charge[0], rtol=1e-10)
    rays, _ = _run(detector, medium, elements, alpha, force_rays=True, rel_tol=1e-9)
    np.testing.assert_allclose(rays, charge, rtol=1e-8)


@pytest.fixture(scope="module")
def demo():
    from lighthit_amd.synthetic import plain_small
    w = plain_small()
    optics = prepare_optics(w.medium, w.detector)
    alpha = acceptance_alpha(w.detector)
    return w, optics, alpha


def test_cap_and_direct_ray_evaluators_agree_on_demo_event(demo):
    w, optics, alpha = demo
    d = w.detector
    radii = d.receiver_surface.radii(len(d))
    # a subset of OMs to keep the direct-ray evaluation short
    ids = np.array([4, 22, 35])
    sub = d.subset(ids)
    elements = w.elements.subset(np.arange(len(w.elements)) % 4 == 1)
    cap, _ = order0.first_entry(sub.positions_m, (-sub.orientations), sub.effective_area_m2, radii[ids],
                                elements, optics, alpha, rel_tol=1e-10)
    rays, _ = order0.first_entry(sub.positions_m, (-sub.orientations), sub.effective_area_m2, radii[ids],
                                 elements, optics, alpha, rel_tol=1e-9, force_rays=True)
    assert np.all(cap > 0)
    np.testing.assert_allclose(rays, cap, rtol=1e-8)


def test_splitting_elements_leaves_charge_unchanged(demo):
    w, optics, alpha = demo
    d = w.detector.subset(np.array([10, 29]))
    e = w.elements.subset(np.arange(len(w.elements)) % 4 == 2)
    half = e.length_m / 2
    tmid = 0.5 * (e.start_ns + e.end_ns)
    split = replace(e, start_m=np.vstack([e.start_m, e.start_m + half[:, None] * e.direction]),
                    direction=np.vstack([e.direction, e.direction]), length_m=np.r_[half, half],
                    coefficient0=np.r_[e.coefficient0, e.coefficient0] / 2,
                    coefficient2=np.r_[e.coefficient2, e.coefficient2] / 2,
                    beta=np.r_[e.beta, e.beta], start_ns=np.r_[e.start_ns, tmid],
                    end_ns=np.r_[tmid, e.end_ns], row_index=np.r_[e.row_index, e.row_index],
                    uid=np.r_[e.uid, e.uid])
    radii = d.receiver_surface.radii(len(d))
    q1, _ = order0.first_entry(d.positions_m, (-d.orientations), d.effective_area_m2, radii, e, optics, alpha)
    q2, _ = order0.first_entry(d.positions_m, (-d.orientations), d.effective_area_m2, radii, split, optics, alpha)
    np.testing.assert_allclose(q2, q1, rtol=1e-9)


def test_bins_hold_the_charge_inside_a_wide_window(demo):
    w, optics, alpha = demo
    d = w.detector.subset(np.array([16]))
    radii = d.receiver_surface.radii(len(d))
    charge, bins = order0.first_entry(d.positions_m, (-d.orientations), d.effective_area_m2, radii, w.elements,
                                      optics, alpha, origins=np.zeros(1),
                                      edges=np.arange(-100., 2000., 5.))
    np.testing.assert_allclose(bins.sum(axis=1), charge, rtol=1e-10)
    assert np.all(charge > 0)


def test_leading_model_reproduces_the_published_demonstration_total(demo):
    """README.md:133 quotes 63.74 pe of order 0 for the demonstration shower: the
    value of the leading finite-cone model (ballistic_interval.py:163-265), which
    the oracle restates independently of its first-entry evaluator."""
    w, optics, alpha = demo
    d = w.detector
    leading = order0.leading_interval_charge(d.positions_m, (-d.orientations), d.effective_area_m2,
                                             d.receiver_surface.radii(len(d)), w.elements,
                                             optics, alpha)
    assert abs(leading.sum() - 63.74) <= 0.005


def test_leading_model_axial_limit_is_the_point_receiver_integral():
    """Far from a small sphere, the window law integrates to one: the leading model
    equals the point-receiver line integral of _integral (closed form in lambda for
    a constant-phase medium: one Gauss node is exact)."""
    beta, length = 1.0, 0.01
    medium = SpectralMedium([400., 500.], [.02, .02], [.01, .01], [1.34, 1.34], [1.37, 1.37], g=.9)
    rho = 40.0
    n = 1.34
    sine = np.sqrt(1 - 1 / (beta * n) ** 2)
    cosine = 1 / (beta * n)
    # OM on the cone of the element's midpoint
    centre = np.array([rho, 0., length / 2 + rho * cosine / sine])
    look = np.array([0., 0., 1.])
    detector = DetectorArray([centre], [-look], .05, lambda x: .45 + .45 * x + .1 * x * x,
                             lambda lam: np.full_like(lam, .25),
                             receiver_surface=SphereSurface(1e-4))
    alpha = acceptance_alpha(detector)
    elements = _element(length, beta)
    optics = prepare_optics(medium, detector)
    optics = dict(optics, phase_nodes=optics["phase_nodes"] + 1e-12 * np.arange(
        len(optics["phase_nodes"])))
    got = order0.leading_interval_charge(detector.positions_m, (-detector.orientations), .05, 1e-4,
                                         elements, optics, alpha, rel_tol=1e-10)
    # photons per nm per m: C0/lambda^2 (1 - 1/(beta n)^2); geometry 1/(2 pi rho sin)
    lam = np.linspace(400., 500., 20001)
    spectral = np.trapezoid(.25 * 1e8 * length * sine ** 2 / lam ** 2, lam)
    distance = rho / sine
    accept = (.45 + .45 * cosine + .1 * cosine ** 2)
    expected = spectral / length * np.exp(-.03 * distance) * accept / (2 * np.pi * rho * sine) * .05
    np.testing.assert_allclose(got, [expected], rtol=1e-6)
