"""CPU pins of the stage (b) oracle (Wigner d, harmonics, rotation, eq. (*))."""
from math import factorial

import numpy as np
import pytest
from scipy.special import sph_harm_y

from oracle import ge2 as oracle
from ge2_cases import make_case


def wigner_formula(j, mp, m, beta):
    """Wigner's explicit sum for d^j_(m' m)(beta)."""
    total = 0.0
    c, s = np.cos(beta / 2), np.sin(beta / 2)
    for k in range(0, 2 * j + 1):
        a, b, d, e = j + m - k, k, j - k - mp, k - m + mp
        if min(a, b, d, e) < 0:
            continue
        total += ((-1) ** (k - m + mp) * np.sqrt(float(factorial(j + mp) * factorial(j - mp)
                                                       * factorial(j + m) * factorial(j - m)))
                  / (factorial(a) * factorial(b) * factorial(d) * factorial(e))
                  * c ** (2 * j - 2 * k + m - mp) * s ** (2 * k - m + mp))
    return total


def real_harmonics(degree, vector):
    """Real orthonormal harmonics from SciPy's complex ones (no Condon-Shortley phase)."""
    x, y, z = vector / np.linalg.norm(vector)
    theta, phi = np.arccos(np.clip(z, -1, 1)), np.arctan2(y, x)
    out = np.zeros((degree + 1) ** 2)
    for l in range(degree + 1):
        for m in range(-l, l + 1):
            value = sph_harm_y(l, abs(m), theta, phi) * (-1) ** abs(m)
            if m == 0:
                out[l * l + l] = value.real
            elif m > 0:
                out[l * l + l + m] = np.sqrt(2) * value.real
            else:
                out[l * l + l + m] = np.sqrt(2) * value.imag
    return out


@pytest.mark.parametrize("cosine", [1.0, 0.93, 0.1, -0.4, -1.0])
def test_wigner_matches_explicit_formula(cosine):
    degree, max_out, max_in = 9, 3, 4
    got = oracle.wigner(degree, max_out, max_in, cosine)
    beta = np.arccos(cosine)
    for l in range(degree + 1):
        for nu in range(-max_out, max_out + 1):
            for mu in range(max_in + 1):
                expected = wigner_formula(l, mu, nu, beta) if abs(nu) <= l and mu <= l else 0.0
                assert got[l, nu + max_out, mu] == pytest.approx(expected, abs=1e-12)


def test_harmonics_match_scipy():
    rng = np.random.default_rng(1)
    for _ in range(20):
        v = rng.normal(size=3)
        np.testing.assert_allclose(oracle.harmonics(v, 5), real_harmonics(5, v), atol=1e-12)
    np.testing.assert_allclose(oracle.harmonics([0, 0, 2.0], 4), real_harmonics(4, np.array([0, 0, 1.0])),
                               atol=1e-12)


def _frame_of(cosine, phi):
    sine = np.sqrt(1 - cosine ** 2)
    e1 = np.array([cosine * np.cos(phi), cosine * np.sin(phi), -sine])
    e2 = np.array([-np.sin(phi), np.cos(phi), 0.0])
    e3 = np.array([sine * np.cos(phi), sine * np.sin(phi), cosine])
    return np.stack((e1, e2, e3))


@pytest.mark.parametrize("cosine,phi", [(0.3, 1.1), (-0.8, -2.5), (0.999, 0.2), (1.0, 0.0)])
def test_rotation_rows_rotate_real_harmonics(cosine, phi):
    # sum_nu E_(nu mu) Y_(l nu)(x') = Y_(l mu)(x) with x' = x in the r-hat frame.
    degree, max_m = 6, 3
    rows = oracle.rotation_rows(degree, degree, max_m, cosine, phi)   # max_out = degree: complete
    frame = _frame_of(cosine, phi)
    rng = np.random.default_rng(7)
    for _ in range(5):
        x = rng.normal(size=3)
        x /= np.linalg.norm(x)
        y, yp = real_harmonics(degree, x), real_harmonics(degree, frame @ x)
        for l in range(degree + 1):
            for mu in range(-min(l, max_m), min(l, max_m) + 1):
                lhs = sum(rows[l, degree + nu, max_m + mu] * yp[l * l + l + nu]
                          for nu in range(-l, l + 1))
                assert lhs == pytest.approx(y[l * l + l + mu], abs=1e-11)


def _rotation_by_fit(l, frame, rng):
    """Real rotation D with Y(x) = Y(frame x) @ D, from 4l+8 sample directions."""
    x = rng.normal(size=(4 * l + 8, 3))
    x /= np.linalg.norm(x, axis=1)[:, None]
    index = slice(l * l, (l + 1) ** 2)
    y = np.array([real_harmonics(l, v)[index] for v in x])
    yp = np.array([real_harmonics(l, frame @ v)[index] for v in x])
    return np.linalg.lstsq(yp, y, rcond=None)[0]   # (nu, mu)


def dense_equation_star(bundle, sources, positions, looks, alpha, area, window):
    """Equation (*) (directional.py:17-37) with SciPy rotations; (frequency, receiver)."""
    rng = np.random.default_rng(11)
    omega = bundle["omega"]
    degree, A = bundle["degree"], bundle["acceptance_degree"]
    coef = bundle["coefficients"].astype(np.complex128)
    log_r = bundle["log_r"]
    frame = sources[0].frame
    points = frame.rotate(sources[0].points_m() - frame.centre_m)
    receivers = frame.rotate(positions - frame.centre_m)
    lk = frame.rotate(looks)
    M = sources[0].azimuthal_degree
    out = np.zeros((len(omega), len(positions)), complex)
    for d in range(len(positions)):
        for p in range(len(points)):
            v = receivers[d] - points[p]
            r = np.linalg.norm(v)
            if r < window[0] or r > window[1]:
                continue
            cosine, phi = np.clip(v[2] / r, -1, 1), np.arctan2(v[1], v[0])
            fr = _frame_of(cosine, phi)
            yd = real_harmonics(A, fr @ lk[d])
            j = int(np.clip(np.searchsorted(log_r, np.log(r), side="right") - 1, 0, len(log_r) - 2))
            t = np.log(r) - log_r[j]
            for lam, src in enumerate(sources):
                phase = np.exp(1j * omega * src.reference_ns)
                scale = area[d] * np.exp(-bundle["absorption"][lam] * r) / (4 * np.pi * r * r)
                flight = np.exp(1j * omega * r / bundle["speed"][lam])
                cursor = 0
                for l in range(sources[0].degree + 1):
                    width = 2 * min(l, M) + 1
                    if l > degree:
                        break
                    D = _rotation_by_fit(l, fr, rng) if l else np.ones((1, 1))
                    full = np.zeros((2 * l + 1, len(omega)), complex)
                    keep = min(l, M)
                    full[l - keep:l + keep + 1] = src.channels[p, cursor:cursor + width]
                    rotated = D @ full                     # (nu, w) in the r-hat frame
                    cursor += width
                    for k, (a, mu) in enumerate(zip(bundle["pair_lambda"], bundle["pair_mu"])):
                        if mu > l:
                            continue
                        c = coef[lam, j, l, k]
                        radial = ((c[0] * t + c[1]) * t + c[2]) * t + c[3]
                        for sign in ((1,) if mu == 0 else (1, -1)):
                            nu = sign * mu
                            out[:, d] += (alpha[a] * yd[a * a + a + nu] * radial
                                          * rotated[l + nu] * scale * flight * phase)
    return out


def test_apply_matches_dense_equation_star():
    bundle, sources, positions, looks, alpha, area = make_case(
        5, bands=2, degree=4, azimuthal=2, acceptance=2, freq=3, points=6, receivers=4,
        intervals=12)
    window = (2.0, 150.0)
    got = oracle.apply_multi_spectral(bundle, sources, positions, looks, alpha,
                                      np.repeat(area[None], 2, axis=0), radial_window_m=window)
    expected = dense_equation_star(bundle, sources, positions, looks, alpha, area, window)
    # The oracle rounds the band-aligned channels to complex64 as the reference does.
    scale = np.abs(expected).max()
    np.testing.assert_allclose(got, expected, rtol=0, atol=1e-6 * scale)


def test_apply_is_linear_in_channels():
    bundle, sources, positions, looks, alpha, area = make_case(2, freq=5, points=8, receivers=5)
    from dataclasses import replace
    window = (2.0, 150.0)
    full = oracle.apply_folded(bundle, sources, positions, looks, alpha, area, window)
    half = [replace(s, channels=s.channels * 0.5) for s in sources]
    np.testing.assert_allclose(oracle.apply_folded(bundle, half, positions, looks, alpha, area, window),
                               0.5 * full, rtol=1e-6, atol=1e-9 * np.abs(full).max())


def test_radial_window_rejects_outside_cache():
    bundle, sources, positions, looks, alpha, area = make_case(3, freq=2, points=4, receivers=2)
    with pytest.raises(ValueError):
        oracle.apply_folded(bundle, sources, positions, looks, alpha, area, (1.0, 150.0))
