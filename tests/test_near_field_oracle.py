"""CPU pins of the near-field clamp oracle (order >=2 of sources inside the
tables' radial minimum)."""
import numpy as np
import pytest

from oracle import ge2 as oracle_ge2
from oracle import near_field as oracle_near


def _bundle(bands=None):
    from lighthit_amd.fold import KernelPack
    from lighthit_amd.route import RouteConfig, RouteKernel
    from lighthit_amd.synthetic import plain_small, synthetic_pack
    w = plain_small()
    kernel = RouteKernel(w.medium, w.detector, RouteConfig(relative_time_edges_ns=w.edges_ns))
    pack = KernelPack()
    pack.add(*synthetic_pack(w.medium, kernel.wavelength))
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, 8, np.arange(0, 65, 8))
    if bands is not None:
        bundle = dict(bundle, coefficients=bundle["coefficients"][bands],
                      absorption=np.asarray(bundle["absorption"])[bands],
                      speed=np.asarray(bundle["speed"])[bands])
    return w, kernel, bundle


def _evaluate(bundle, impact, u, look, harmonics, weights, alpha, clamp_low):
    low, high = np.exp(bundle["log_r"][[0, -1]])
    return oracle_near.cone_many_looks(
        impact, u, look, harmonics, weights, alpha, bundle["pair_lambda"], bundle["pair_mu"],
        bundle["log_r"], bundle["coefficients"], bundle["absorption"], bundle["speed"],
        bundle["omega"], bundle["phase"] == "flight", low, high, clamp_low)[0]


def test_cone_evaluator_equals_the_general_apply_outside_the_minimum():
    """An axially symmetric point source: the m = 0 Wigner column against the
    explicit rotation rows of the general order >=2 apply."""
    w, kernel, bundle = _bundle()
    rng = np.random.default_rng(3)
    degree, bands = 8, len(bundle["absorption"])
    harmonics = rng.normal(size=(bands, degree + 1))
    weights = rng.uniform(0.5, 2.0, bands)
    alpha = np.asarray(kernel.acceptance()["alpha"], float)
    look = rng.normal(size=3)
    look /= np.linalg.norm(look)
    impact, u = 5.0, np.linspace(-30.0, 40.0, 9)
    got = _evaluate(bundle, impact, u, look, harmonics, weights, alpha, 2.0)
    receivers = np.column_stack([np.full(len(u), impact), np.zeros(len(u)), u])
    channels = np.broadcast_to(harmonics[:, None, :, None],
                               (bands, 1, degree + 1, len(bundle["omega"]))).astype(np.complex64)
    want = oracle_ge2.apply_core(
        np.zeros((1, 3)), receivers, np.tile(look, (len(u), 1)), channels,
        np.repeat(weights[:, None], len(u), axis=1), oracle_ge2._starts(degree, 0), degree, 0,
        int(bundle["acceptance_degree"]), alpha, bundle["pair_lambda"], bundle["pair_mu"],
        bundle["log_r"], bundle["coefficients"], bundle["absorption"], bundle["speed"],
        bundle["omega"], True, 2.0, 150.0)
    scale = np.abs(want).max()
    assert scale > 0
    # the general apply takes complex64 channels
    assert np.abs(got - want).max() < 2e-6 * scale


def test_clamp_keeps_the_minimum_amplitude_with_the_true_flight_phase():
    w, kernel, bundle = _bundle(bands=[1])
    rng = np.random.default_rng(4)
    harmonics = rng.normal(size=(1, 9))
    alpha = np.asarray(kernel.acceptance()["alpha"], float)
    look = np.array([0.3, -0.5, 0.81])
    look /= np.linalg.norm(look)
    low = float(np.exp(bundle["log_r"][0]))
    omega = np.asarray(bundle["omega"])
    speed = float(bundle["speed"][0])
    for impact, along in ((0.5, 0.7), (1.2, -0.3), (0.0001, 1.5)):
        radius = np.hypot(impact, along)
        assert radius < low
        near = _evaluate(bundle, impact, [along], look, harmonics, [1.0], alpha, 0.0)[0]
        at_low = _evaluate(bundle, impact * low / radius, [along * low / radius], look,
                           harmonics, [1.0], alpha, 0.0)[0]
        want = at_low * np.exp(1j * omega * (radius - low) / speed)
        np.testing.assert_allclose(near, want, rtol=1e-12, atol=1e-15 * np.abs(want).max())
        # without the clamp the point is left out
        off = _evaluate(bundle, impact, [along], look, harmonics, [1.0], alpha, low)[0]
        assert not np.any(off)


def test_near_portion_is_the_chord_of_the_sphere():
    from types import SimpleNamespace
    start = np.array([[-5.0, 0.7, 0.0], [-5.0, 3.0, 0.0], [-1.0, 0.7, 0.0]])
    elements = SimpleNamespace(start_m=start, direction=np.tile([1.0, 0.0, 0.0], (3, 1)),
                               length_m=np.array([10.0, 10.0, 0.5]))
    portions = oracle_near.near_portions(np.zeros((1, 3)), elements, 2.0)
    half = np.sqrt(4.0 - 0.49)
    assert list(portions) == [0]
    (i0, a0, b0), (i2, a2, b2) = portions[0]
    assert i0 == 0 and abs(a0 - (5.0 - half)) < 1e-12 and abs(b0 - (5.0 + half)) < 1e-12
    assert i2 == 2 and a2 == 0.0 and b2 == 0.5
