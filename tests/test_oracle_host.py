"""CPU pins of the oracle's host restatements (oracle/host.py) and their agreement
with the package's host code."""
from dataclasses import replace

import numpy as np
import pytest

from oracle import host


class _Field:
    def __init__(self, start, direction, length, photons, t0=None):
        self.start_m = np.asarray(start, float)
        self.direction = np.asarray(direction, float)
        self.length_m = np.asarray(length, float)
        self.photons = np.asarray(photons, float)
        self.start_ns = np.zeros(len(self.length_m)) if t0 is None else np.asarray(t0, float)
        self.end_ns = self.start_ns + 1.0

    @property
    def midpoints_m(self):
        return self.start_m + 0.5 * self.length_m[:, None] * self.direction


def test_frame_of_a_line_source_is_the_line():
    axis = np.array([0.3, -0.4, 0.866])
    axis /= np.linalg.norm(axis)
    s = np.linspace(-5, 5, 41)
    rng = np.random.default_rng(1)
    start = 10 + s[:, None] * axis
    direction = axis + 0.01 * rng.normal(size=(41, 3))
    direction /= np.linalg.norm(direction, axis=1)[:, None]
    field = _Field(start, direction, np.full(41, 0.01), rng.uniform(1, 2, 41))
    frame = host.axis_frame(field)
    assert abs(abs(frame.axis @ axis) - 1) < 1e-4
    assert frame.axis @ axis > 0          # along the light
    basis = np.stack([frame.first, frame.second, frame.axis])
    np.testing.assert_allclose(basis @ basis.T, np.eye(3), atol=1e-12)
    share = field.photons / field.photons.sum()
    np.testing.assert_allclose(frame.centre_m, share @ field.midpoints_m, rtol=1e-12)


def test_frame_of_a_point_source_is_its_mean_direction():
    direction = np.array([[0.0, 0.6, 0.8], [0.0, 0.8, 0.6]])
    field = _Field(np.zeros((2, 3)), direction, np.full(2, 1e-12), [1.0, 3.0])
    frame = host.axis_frame(field)
    mean = np.array([0.0, 0.75, 0.65])
    np.testing.assert_allclose(frame.axis, mean / np.linalg.norm(mean), atol=1e-9)


def test_frame_equals_the_package_frame_on_the_demonstration_shower():
    from lighthit_amd.axial import AxisFrame
    from lighthit_amd.synthetic import plain_small
    w = plain_small()
    field = w.elements.field(0, phase_index=1.344)
    a, b = host.axis_frame(field), AxisFrame.of(field)
    for name in ("centre_m", "axis", "first", "second"):
        np.testing.assert_allclose(getattr(a, name), getattr(b, name), rtol=0, atol=1e-12)


def test_optical_panels_hold_the_efficiency_and_match_the_package():
    from lighthit_amd.optics import prepare_optics
    from lighthit_amd.synthetic import plain_small
    w = plain_small()
    mine = host.prepare_optics(w.medium, w.detector)
    edges = mine["panel_edges"]
    assert set(w.medium.wavelength_nm) <= set(edges)
    # every panel holds the efficiency at points between the probes too
    for k, c in enumerate(mine["efficiency_coefficients"]):
        q = np.linspace(0, 1, 33)
        lam = edges[k] + q * (edges[k + 1] - edges[k])
        exact = w.detector.spectral_weight(lam)
        fitted = np.polynomial.polynomial.polyval(q, c)
        assert np.max(np.abs(fitted - exact) / exact) < 1e-7
    theirs = prepare_optics(w.medium, w.detector)
    np.testing.assert_array_equal(edges, theirs["panel_edges"])
    np.testing.assert_allclose(mine["efficiency_coefficients"],
                               theirs["efficiency_coefficients"], rtol=0, atol=1e-13)
    for name in ("phase_nodes", "group_nodes", "extinction_nodes"):
        np.testing.assert_allclose(mine[name], theirs[name], rtol=1e-15)


def _bands_inputs(nodes):
    from lighthit_amd.model import WavelengthQuadrature
    from lighthit_amd.synthetic import plain_small
    w = plain_small()
    quadrature = WavelengthQuadrature.gauss_legendre(400., 500., nodes)
    phase = w.medium.sample(quadrature.wavelength_nm)["phase_index"]
    efficiency = w.detector.spectral_weight(quadrature.wavelength_nm)
    return w, quadrature, phase, efficiency


def test_node_bands_at_full_rank_interpolate_exactly():
    w, q, phase, efficiency = _bands_inputs(3)
    fold = host.node_bands(q.wavelength_nm, q.weight_nm, efficiency, phase, 8, 3)
    assert fold.bands == (0, 1, 2)
    # every column is reproduced (degrees 0 and 1 span only two functions of beta,
    # so their interpolation is a projection, not the identity)
    assert fold.worst_design_residual < 1e-9
    np.testing.assert_allclose(fold.interpolation[2:], np.broadcast_to(np.eye(3), (7, 3, 3)),
                               atol=1e-6)


def test_node_bands_equal_the_package_fold():
    from lighthit_amd.fold import BandFold
    from lighthit_amd.route import RouteConfig, RouteKernel
    w, q, phase, efficiency = _bands_inputs(8)
    kernel = RouteKernel(w.medium, w.detector, RouteConfig(wavelength_nodes=8))
    theirs = BandFold.of(kernel, rank=3)
    # the package's subset is optimal to the rounding of the residuals
    mine = host.node_bands(q.wavelength_nm, q.weight_nm, efficiency, phase, 8, 3,
                           bands=theirs.bands)
    assert 0 < mine.worst_design_residual <= 1
    assert mine.worst_design_residual <= mine.best_design_residual * (1 + 1e-6)
    assert np.isclose(theirs.metadata()["worst_residual_design"], mine.worst_design_residual,
                      rtol=1e-6)
    # a clearly worse subset is not chosen
    other = host.node_bands(q.wavelength_nm, q.weight_nm, efficiency, phase, 8, 3,
                            bands=(0, 1, 7))
    assert other.worst_design_residual > 2 * mine.best_design_residual
    np.testing.assert_allclose(mine.interpolation, theirs.interpolation, rtol=0, atol=1e-6)
    np.testing.assert_allclose(mine.scale, theirs.scale, rtol=1e-15)


def test_segment_bounds_equal_a_brute_force_scan():
    rng = np.random.default_rng(4)
    count = 40
    direction = rng.normal(size=(count, 3))
    direction /= np.linalg.norm(direction, axis=1)[:, None]
    elements = _Field(rng.uniform(-20, 20, (count, 3)), direction, rng.uniform(0.5, 8, count),
                      np.ones(count), rng.uniform(0, 50, count))
    positions = rng.uniform(-40, 40, (13, 3))
    s = np.linspace(0, 1, 4001)
    points = (elements.start_m[None] + (s[:, None] * elements.length_m[None])[..., None]
              * elements.direction[None])                               # (s, element, 3)
    distance = np.linalg.norm(positions[:, None, None] - points[None], axis=-1)  # (om, s, el)
    brute = distance.min(axis=(1, 2))
    mine = host.min_distance(positions, elements)
    assert np.all(mine <= brute + 1e-12)
    np.testing.assert_allclose(mine, brute, atol=2e-3 * 8 / 4000 * 4)
    n_g = 1.37
    brute_arrival = (elements.start_ns[None, None] + n_g / 0.299792458 * distance).min(axis=(1, 2))
    arrival = host.earliest_arrival(positions, elements, n_g)
    assert np.all(arrival <= brute_arrival + 1e-9)
    np.testing.assert_allclose(arrival, brute_arrival, atol=0.05)


def test_delay_groups_of_a_shifted_copy():
    from lighthit_amd.synthetic import plain_small
    first = plain_small().elements
    late = replace(first, start_ns=first.start_ns + 250.0, end_ns=first.end_ns + 250.0)
    names = ("start_m", "direction", "length_m", "coefficient0", "coefficient2", "beta",
             "start_ns", "end_ns", "row_index", "uid")
    both = replace(first, **{name: np.concatenate([getattr(first, name), getattr(late, name)])
                             for name in names})
    prompt, groups, beyond = host.delay_groups(both, prompt_delay_ns=20.0, cell_ns=100.0,
                                               window_end_ns=740.0, group_index_max=1.386)
    half = len(first)
    assert prompt[:half].all() and not prompt[half:].any() and not beyond.any()
    assert len(groups) == 1
    mask, shift = groups[0]
    assert mask[half:].all() and not mask[:half].any()
    assert 249.0 < shift <= 250.0
    from lighthit_amd.route import delay_groups
    p2, g2, b2 = delay_groups(both, prompt_delay_ns=20.0, cell_ns=100.0, window_end_ns=740.0,
                              group_index_max=1.386)
    assert np.array_equal(p2, prompt) and np.array_equal(b2, beyond)
    assert len(g2) == 1 and np.array_equal(g2[0][0], mask) and g2[0][1] == shift


@pytest.mark.parametrize("delay", [0.0, 1e4])
def test_far_delayed_steps_only_add_charge(delay):
    from lighthit_amd.synthetic import plain_small
    first = plain_small().elements
    late = replace(first, start_ns=first.start_ns + delay, end_ns=first.end_ns + delay)
    names = ("start_m", "direction", "length_m", "coefficient0", "coefficient2", "beta",
             "start_ns", "end_ns", "row_index", "uid")
    both = replace(first, **{name: np.concatenate([getattr(first, name), getattr(late, name)])
                             for name in names})
    prompt, groups, beyond = host.delay_groups(both, prompt_delay_ns=20.0, cell_ns=20.0,
                                               window_end_ns=740.0, group_index_max=1.386)
    if delay == 0.0:
        assert prompt.all() and not groups
    else:
        assert beyond[len(first):].all() and not groups
