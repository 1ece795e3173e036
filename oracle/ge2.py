"""Oracle of stage (b): the folded order >=2 apply in FP64.

``apply_multi_spectral`` restates ``_shower_basis.py:1126-1198`` (band
alignment with complex64 rounding of the aligned channels, frame rotation,
radial-window checks, final carrier) around the C restatement of
``_apply_multi_spectral`` (``oracle/ge2_apply.c``); ``apply_folded`` restates
``shower/ge2.py:99-119``.

Parity pins (tests/test_ge2_oracle.py):
  * Wigner small-d against the explicit factorial sum (Wigner's formula);
  * detector harmonics against SciPy's complex spherical harmonics;
  * rotation rows E^l through the rotation identity
    sum_nu E_(nu mu) Y_(l nu)(x') = Y_(l mu)(x), x' = x in the r-hat frame;
  * the whole apply against a dense numpy evaluation of equation (*)
    (``directional.py:17-37``) that builds every rotation from SciPy
    harmonics by least squares instead of the Wigner recurrence.
"""
import ctypes

import numpy as np

from ._clib import arr, lib


def wigner(degree, max_out, max_in, cosine):
    """``d^l_(mu nu)(theta)`` as out[l, nu + max_out, mu] (``_directional_numba.py:52-88``)."""
    out = np.zeros((degree + 1, 2 * max_out + 1, max_in + 1))
    lib().oracle_wigner(ctypes.c_int(degree), ctypes.c_int(max_out), ctypes.c_int(max_in),
                        ctypes.c_double(cosine), out.ctypes.data_as(ctypes.c_void_p))
    return out


def harmonics(vector, degree):
    """Real orthonormal harmonics, flat ``l*l+l+m`` (``_directional_numba.py:91-126``)."""
    out = np.zeros((degree + 1) ** 2)
    x, y, z = map(float, vector)
    lib().oracle_harmonics(ctypes.c_double(x), ctypes.c_double(y), ctypes.c_double(z),
                           ctypes.c_int(degree), out.ctypes.data_as(ctypes.c_void_p))
    return out


def rotation_rows(degree, max_out, max_m, cosine, phi):
    """Dense E^l[a, b] of every degree (``_shower_basis.py:1006-1029``)."""
    out = np.zeros((degree + 1, 2 * max_out + 1, 2 * max_m + 1))
    lib().oracle_rotation_rows(ctypes.c_int(degree), ctypes.c_int(max_out), ctypes.c_int(max_m),
                               ctypes.c_double(cosine), ctypes.c_double(phi),
                               out.ctypes.data_as(ctypes.c_void_p))
    return out


def _starts(degree, azimuthal_degree):
    starts = [0]
    for ell in range(degree + 1):
        starts.append(starts[-1] + 2 * min(ell, azimuthal_degree) + 1)
    return np.asarray(starts, np.int64)


def apply_core(local_points, local_receivers, local_looks, channels, out_scale, starts, degree,
               max_m, max_out, alpha, pair_lambda, pair_mu, log_r, coefficients, absorption,
               velocity, omega, flight, radial_low, radial_high):
    """``_apply_multi_spectral`` (tolerance 0); returns (receiver, frequency) complex128."""
    points, p_points = arr(local_points, np.float64)
    receivers, p_rec = arr(local_receivers, np.float64)
    looks, p_looks = arr(local_looks, np.float64)
    chan, p_chan = arr(channels, np.complex64)
    scale, p_scale = arr(out_scale, np.float64)
    st, p_st = arr(starts, np.int64)
    al, p_al = arr(alpha, np.float64)
    pl, p_pl = arr(pair_lambda, np.int64)
    pm, p_pm = arr(pair_mu, np.int64)
    lr, p_lr = arr(log_r, np.float64)
    coef, p_coef = arr(coefficients, np.complex64)
    ab, p_ab = arr(absorption, np.float64)
    ve, p_ve = arr(velocity, np.float64)
    om, p_om = arr(omega, np.float64)
    nlam, npoints, nchan, nfreq = chan.shape
    if coef.shape[0] != nlam or coef.shape[-1] != nfreq:
        raise ValueError("channels and coefficients disagree in bands or frequencies")
    result = np.zeros((len(receivers), nfreq), np.complex128)
    i, d = ctypes.c_int, ctypes.c_double
    lib().oracle_apply_multi_spectral(
        p_points, i(npoints), p_rec, p_looks, i(len(receivers)), p_chan, i(nchan), p_scale, p_st,
        i(degree), i(max_m), i(max_out), p_al, p_pl, p_pm, i(len(pl)), p_lr, i(len(lr) - 1),
        p_coef, p_ab, p_ve, p_om, i(nfreq), i(nlam), i(int(bool(flight))), d(radial_low),
        d(radial_high), result.ctypes.data_as(ctypes.c_void_p))
    return result


def apply_multi_spectral(bundle, sources, receivers_m, look_directions, alpha, out_scale, *,
                         radial_window_m=None, channel_dtype=np.complex64):
    """``_shower_basis.py:1126-1198``: (frequency, receiver) complex128 in absolute time."""
    sources = list(sources)
    if not sources or len(sources) != len(bundle["absorption"]):
        raise ValueError("one spectral source is needed for each bundled wavelength")
    first = sources[0]
    omega = np.asarray(bundle["omega"])
    points = np.asarray(first.points_m(), float)
    for item in sources:
        if (not np.array_equal(item.points_m(), points)
                or item.channels.shape != first.channels.shape
                or item.azimuthal_degree != first.azimuthal_degree
                or not np.array_equal(item.frame.axis, first.frame.axis)):
            raise ValueError("fused spectral sources must share points and angular layout")
    reference = min(float(item.reference_ns) for item in sources)
    channels = np.empty((len(sources),) + first.channels.shape, channel_dtype)
    for index, item in enumerate(sources):
        phase = np.exp(1j * omega * (item.reference_ns - reference))
        channels[index] = item.channels * phase[None, None, :]
    starts = _starts(first.degree, first.azimuthal_degree)
    centre = np.asarray(first.frame.centre_m, float)
    local_points = first.frame.rotate(points - centre)
    receivers = np.atleast_2d(np.asarray(receivers_m, float))
    local_receivers = first.frame.rotate(receivers - centre)
    local_looks = first.frame.rotate(np.atleast_2d(np.asarray(look_directions, float)))
    radial_low = float(np.exp(bundle["log_r"][0]))
    radial_high = float(np.exp(bundle["log_r"][-1]))
    if radial_window_m is None:
        for receiver in local_receivers:
            radius = np.linalg.norm(local_points - receiver, axis=1)
            if np.any(radius < radial_low) or np.any(radius > radial_high):
                raise ValueError("Source/receiver distances outside the cache")
    else:
        low, high = map(float, radial_window_m)
        if (not np.isfinite(low) or not np.isfinite(high) or low < radial_low - 1e-9
                or high > radial_high + 1e-9 or low >= high):
            raise ValueError("radial window must lie inside the kernel cache")
        radial_low, radial_high = low, high
    acceptance_degree = int(bundle["acceptance_degree"])
    alpha = np.asarray(alpha, float)[:acceptance_degree + 1]
    max_m = min(int(first.azimuthal_degree), int(bundle["degree"]))
    result = apply_core(local_points, local_receivers, local_looks, channels, out_scale, starts,
                        int(bundle["degree"]), max_m, acceptance_degree, alpha,
                        bundle["pair_lambda"], bundle["pair_mu"], bundle["log_r"],
                        bundle["coefficients"], bundle["absorption"], bundle["speed"], omega,
                        bundle["phase"] == "flight", radial_low, radial_high)
    return result.T * np.exp(1j * omega * reference)[:, None]


def apply_folded(folded, sources, positions, looks, alpha, area, window):
    """``shower/ge2.py:99-119``: node-band sources, detector scale = OM area."""
    if not len(positions):
        return np.zeros((len(folded["omega"]), 0), complex)
    out_scale = np.repeat(np.asarray(area, float)[None, :], len(sources), axis=0)
    points = sources[0].points_m()
    if all(np.array_equal(item.points_m(), points) for item in sources[1:]):
        return apply_multi_spectral(folded, sources, positions, looks, alpha, out_scale,
                                    radial_window_m=window)
    total = 0
    for j, item in enumerate(sources):
        band = dict(folded, coefficients=folded["coefficients"][j:j + 1],
                    absorption=folded["absorption"][j:j + 1], speed=folded["speed"][j:j + 1])
        total = total + apply_multi_spectral(band, [item], positions, looks, alpha,
                                             out_scale[j:j + 1], radial_window_m=window)
    return total
