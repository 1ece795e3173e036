"""Oracle of stage (c): FP64 readout of order >=2 spectra into time bins.

Restates ``readout.py:9-40`` (``frequency_weights``, ``inverse_bins``) and the
readout step of ``shower/api.py:461-467``: demodulate by the OM time origin,
SciPy's not-a-knot ``CubicSpline`` of the real and imaginary parts from the
table grid onto the readout grid (the same third-party call as the reference;
SciPy 1.15.3 here, reference requires ``scipy>=1.11``), then the exact per-bin
Fourier integral.

Parity pins (tests/test_readout_oracle.py):
  * ``inverse_bins`` against the Gaussian-smeared delta of
    ``tests/test_readout.py:10-16`` (closed form through ``ndtr``, 5e-14) and
    the exponential pulse of ``tests/test_readout.py:55-84``;
  * the spline step reproduces any cubic exactly (not-a-knot is exact on
    cubics), and the whole readout of ``exp(i w t0)`` lands in the bin of t0.
"""
import numpy as np


def frequency_weights(omega):
    """``readout.py:9-15``: trapezoid weights of a grid that starts at zero."""
    omega = np.asarray(omega, float)
    if (omega.ndim != 1 or len(omega) < 2 or not np.isfinite(omega).all()
            or omega[0] != 0 or np.any(np.diff(omega) <= 0)):
        raise ValueError("Time inversion requires at least two increasing frequencies starting at zero")
    step = np.diff(omega)
    return np.r_[step[0] / 2, (step[:-1] + step[1:]) / 2, step[-1] / 2]


def inverse_bins(omega, spectrum, edges_ns, sigma_ns=0.0):
    """``readout.py:18-40``: Re of the exact per-bin half-axis Fourier integral; (obs, bin)."""
    omega = np.asarray(omega, float)
    edges = np.asarray(edges_ns, float)
    spectrum = np.asarray(spectrum, complex)
    if edges.ndim != 1 or len(edges) < 2 or not np.isfinite(edges).all() or np.any(np.diff(edges) <= 0):
        raise ValueError("Time edges must be finite and strictly increasing")
    if not np.isfinite(sigma_ns) or sigma_ns < 0:
        raise ValueError("sigma_ns must be finite and nonnegative")
    if spectrum.ndim != 2 or spectrum.shape[0] != len(omega) or not np.isfinite(spectrum).all():
        raise ValueError("spectrum must have shape (frequency, observation)")
    w = frequency_weights(omega)
    out = np.zeros((spectrum.shape[1], len(edges) - 1))
    # One bin at a time, as a plain sum over frequencies.
    for b in range(len(edges) - 1):
        width = edges[b + 1] - edges[b]
        centre = 0.5 * (edges[b] + edges[b + 1])
        row = (width * np.sinc(width * omega / (2 * np.pi)) * np.exp(-1j * centre * omega)
               * w * np.exp(-0.5 * (omega * sigma_ns) ** 2) / np.pi)
        out[:, b] = (row @ spectrum).real
    return out


def route_readout(omega, spectrum, origins, readout_omega, edges_ns):
    """``shower/api.py:461-467``: demodulate, spline onto the readout grid, integrate bins."""
    from scipy.interpolate import CubicSpline
    omega = np.asarray(omega, float)
    readout = np.asarray(readout_omega, float)
    relative = np.asarray(spectrum, complex) * np.exp(
        -1j * omega[:, None] * np.asarray(origins, float)[None, :])
    if len(omega) != len(readout):
        relative = (CubicSpline(omega, relative.real, axis=0)(readout)
                    + 1j * CubicSpline(omega, relative.imag, axis=0)(readout))
    return inverse_bins(readout, relative, edges_ns)
