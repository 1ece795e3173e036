"""CPU pins of the stage (c) oracle and of the host-side readout basis."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.special import ndtr

from oracle import readout as oracle
from lighthit_amd import readout as product


def test_oracle_gaussian_delta_closed_form():
    # tests/test_readout.py:10-16 of the reference.
    omega = np.linspace(0, 5, 2001)
    t0, sigma = 12., 2.
    edges = np.arange(-4, 30, 1.)
    spectrum = np.exp(1j * omega * t0)[:, None]
    actual = oracle.inverse_bins(omega, spectrum, edges, sigma)[0]
    expected = ndtr((edges[1:] - t0) / sigma) - ndtr((edges[:-1] - t0) / sigma)
    np.testing.assert_allclose(actual, expected, atol=5e-14)


def _exponential_pulse(tau_ns, omega, edges):
    spectrum = (1 / (1 - 1j * np.asarray(omega) * tau_ns))[:, None]
    cumulative = np.where(edges > 0, 1 - np.exp(-np.clip(edges, 0, None) / tau_ns), 0.)
    return spectrum, np.diff(cumulative)


def test_oracle_exponential_pulse_on_readout_grid():
    # tests/test_readout.py:73-84 of the reference: spline 65 -> 321 keeps the charge.
    omega, edges = np.linspace(0., 1.2, 65), np.arange(-60., 745., 5.)
    readout = np.linspace(0., 1.2, 321)
    spectrum, exact = _exponential_pulse(20., omega, edges)
    bins = oracle.route_readout(omega, spectrum, np.zeros(1), readout, edges)[0]
    assert bins.sum() == pytest.approx(1., abs=1e-2)
    direct = oracle.inverse_bins(readout, _exponential_pulse(20., readout, edges)[0], edges)[0]
    assert np.abs(bins - direct).sum() < 1e-2
    assert np.abs(bins - exact).sum() < .05


def test_oracle_demodulation_is_a_shift_of_the_clock():
    # exp(i w (t0 + origin)) read out relative to origin peaks in the bin of t0.
    omega = np.linspace(0., 1.2, 65)
    readout = np.linspace(0., 1.2, 321)
    edges = np.arange(-60., 745., 5.)
    origin = np.array([1234.5])
    spectrum = np.exp(1j * omega * (102.5 + origin[0]))[:, None]
    bins = oracle.route_readout(omega, spectrum, origin, readout, edges)[0]
    assert np.argmax(bins) == np.searchsorted(edges, 102.5, side="right") - 1


@pytest.mark.parametrize("invalid", [dict(omega=[0], edges=[0, 1], sigma=0),
                                     dict(omega=[.1, .2], edges=[0, 1], sigma=0),
                                     dict(omega=[0, .1], edges=[1, 0], sigma=0),
                                     dict(omega=[0, .1], edges=[0, 1], sigma=-1)])
def test_invalid_arguments_raise(invalid):
    # tests/test_readout.py:19-21 of the reference, oracle and product host checks.
    omega = np.asarray(invalid["omega"], float)
    spectrum = np.ones((len(omega), 1))
    with pytest.raises(ValueError):
        oracle.inverse_bins(omega, spectrum, invalid["edges"], invalid["sigma"])
    with pytest.raises(ValueError):
        product.bin_kernel(omega, invalid["edges"], invalid["sigma"])


def test_spline_matrix_matches_scipy_not_a_knot():
    rng = np.random.default_rng(0)
    x = np.linspace(0., 1.2, 65)
    at = np.linspace(0., 1.2, 321)
    y = rng.normal(size=(65, 4))
    expected = CubicSpline(x, y, axis=0)(at)
    np.testing.assert_allclose(product.spline_matrix(x, at) @ y, expected, rtol=0, atol=1e-12)
    # nonuniform knots too
    x = np.sort(rng.uniform(0, 3, 12))
    at = np.linspace(x[0], x[-1], 50)
    y = rng.normal(size=12)
    np.testing.assert_allclose(product.spline_matrix(x, at) @ y, CubicSpline(x, y)(at),
                               rtol=0, atol=1e-11)


def test_spline_matrix_exact_on_cubics():
    x = np.linspace(0., 1.2, 65)
    at = np.linspace(0., 1.2, 321)
    y = 2 - x + 3 * x ** 2 - 0.7 * x ** 3
    np.testing.assert_allclose(product.spline_matrix(x, at) @ y,
                               2 - at + 3 * at ** 2 - 0.7 * at ** 3, atol=1e-12)


def test_folded_basis_equals_spline_then_integral():
    # M = K S reproduces the oracle's per-OM spline-then-integrate readout in FP64.
    rng = np.random.default_rng(3)
    omega = np.linspace(0., 1.2, 65)
    readout = np.linspace(0., 1.2, 321)
    edges = np.arange(-60., 745., 5.)
    spectrum = rng.normal(size=(65, 5)) + 1j * rng.normal(size=(65, 5))
    matrix = product.bin_kernel(readout, edges) @ product.spline_matrix(omega, readout)
    expected = oracle.route_readout(omega, spectrum, np.zeros(5), readout, edges)
    np.testing.assert_allclose((matrix @ spectrum).real.T, expected, rtol=0, atol=1e-12)


def test_readout_basis_has_the_same_bits_at_any_thread_count():
    """The folded basis M = K S feeds every process of the served form: it must
    not depend on how many BLAS threads a process runs (a threaded matrix
    product rounds differently), or the same event would get different bins
    alone and served."""
    from threadpoolctl import threadpool_limits
    omega, edges = np.linspace(0., 1.2, 65), np.arange(-60., 745., 1.)
    readout = np.linspace(0., 1.2, 321)
    with threadpool_limits(limits=1):
        one = product.ReadoutBasis(omega, edges, readout, device="cpu")
    with threadpool_limits(limits=8):
        many = product.ReadoutBasis(omega, edges, readout, device="cpu")
    assert np.array_equal(one.matrix, many.matrix)
    # and it is the matrix product
    direct = product.bin_kernel(readout, edges) @ product.spline_matrix(omega, readout)
    np.testing.assert_allclose(one.matrix, direct, rtol=0, atol=1e-14 * np.abs(direct).max())
