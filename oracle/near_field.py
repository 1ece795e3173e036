"""Oracle of the near-field order >=2 clamp in FP64 — TEST INFRASTRUCTURE ONLY.

Restates ``near_field.py:44-173`` (``near_portions``, ``_harmonics``,
``near_field_ge2``) and the fused-looks cone evaluator ``_cone_many_looks``
(``muon/track_ge2_primitive.py:612-660``) it calls through
``cone_point_evaluator`` (``:84-162``): element portions closer to an OM than
the tables' radial minimum get the order >=2 amplitude at that minimum along
the same directions, with the flight phase of the actual distance.

Pins (tests/test_near_field_oracle.py): outside the radial minimum the cone
evaluator equals the general order >=2 apply of ``oracle.ge2`` for an
axially symmetric point source (an independent code path: explicit rotation
rows instead of the m = 0 Wigner column); inside it, every amplitude equals
the one at the minimum with the phase of the true distance; the near portion
of a straight element is the chord of the sphere of the radial minimum.
"""
from __future__ import annotations

import numpy as np

from . import ge2 as oracle_ge2


def near_portions(positions, elements, radius):
    """{om: [(element, s0, s1), ...]}: element portions within ``radius`` of each
    OM, ``s`` along the element from its start (``near_field.py:44-69``)."""
    from scipy.spatial import cKDTree
    positions = np.asarray(positions, float)
    start = np.asarray(elements.start_m, float)
    direction = np.asarray(elements.direction, float)
    length = np.asarray(elements.length_m, float)
    middle = start + 0.5 * length[:, None] * direction
    tree = cKDTree(positions)
    nearest, _ = tree.query(middle, k=1)
    portions = {}
    for index in np.flatnonzero(nearest < radius + 0.5 * length):
        reach = radius + 0.5 * length[index]
        for om in sorted(tree.query_ball_point(middle[index], reach)):
            delta = positions[om] - start[index]
            z = float(delta @ direction[index])
            rho2 = max(float(delta @ delta) - z * z, 0.0)
            if rho2 >= radius * radius:
                continue
            half = np.sqrt(radius * radius - rho2)
            s0, s1 = max(z - half, 0.0), min(z + half, float(length[index]))
            if s1 > s0:
                portions.setdefault(int(om), []).append((int(index), s0, s1))
    return portions


def cone_harmonics(elements, index, wavelengths, phases, degree):
    """Spectral cone harmonics per unit C0 of one element (``near_field.py:72-88``)."""
    from scipy.special import eval_legendre
    ell = np.arange(degree + 1)
    norm = np.sqrt((2 * ell + 1) / (4 * np.pi))
    beta = float(elements.beta[index])
    c0 = float(elements.coefficient0[index])
    ratio = float(elements.coefficient2[index]) / c0 if c0 > 0 else 0.0
    out = np.zeros((len(wavelengths), degree + 1))
    for band, (wavelength, phase) in enumerate(zip(wavelengths, phases)):
        yield_ = (1 + ratio / phase ** 2) / wavelength ** 2
        if beta * phase <= 1 or yield_ <= 0:
            continue
        out[band] = norm * eval_legendre(ell, 1 / (beta * phase)) * yield_
    return out


def cone_many_looks(impact, u, looks, harmonics, weights, alpha, pair_lambda, pair_mu, log_r,
                    coefficients, absorption, velocity, omega, flight, radial_low, radial_high,
                    clamp_low):
    """(looks, points, frequency) order >=2 spectrum of an axial point source at
    the origin seen from (impact, 0, u) (``track_ge2_primitive.py:612-660``)."""
    degree = harmonics.shape[1] - 1
    max_out = int(np.max(pair_lambda))
    looks = np.atleast_2d(np.asarray(looks, float))
    omega = np.asarray(omega, float)
    coefficients = np.asarray(coefficients, complex)
    result = np.zeros((len(looks), len(u), len(omega)), complex)
    for point, along in enumerate(np.asarray(u, float)):
        radius = float(np.hypot(impact, along))
        if radius < clamp_low or radius > radial_high or radius == 0.0:
            continue
        cosine, sine = along / radius, impact / radius
        wigner = oracle_ge2.wigner(degree, max_out, 0, cosine)
        amplitude = max(radius, radial_low)
        j = min(max(int(np.searchsorted(log_r, np.log(amplitude), side="right")) - 1, 0),
                len(log_r) - 2)
        t = np.log(amplitude) - log_r[j]
        scale = (np.asarray(weights, float) * np.exp(-np.asarray(absorption) * amplitude)
                 / (4 * np.pi * amplitude * amplitude))
        angle = (np.outer(1.0 / np.asarray(velocity, float), omega * radius) if flight
                 else np.zeros((len(scale), len(omega))))
        phase = scale[:, None] * np.exp(1j * angle)                  # (bands, freq)
        response = np.zeros((len(pair_lambda), len(omega)), complex)
        for ell in range(degree + 1):
            for p, mu in enumerate(pair_mu):
                if mu > ell:
                    continue
                rotation = (wigner[ell, max_out, 0] if mu == 0
                            else np.sqrt(2.0) * (-1.0) ** mu * wigner[ell, max_out + mu, 0])
                c = coefficients[:, j, ell, p]                      # (bands, 4, freq)
                radial = ((c[:, 0] * t + c[:, 1]) * t + c[:, 2]) * t + c[:, 3]
                response[p] += np.sum(harmonics[:, ell, None] * rotation * radial * phase,
                                      axis=0)
        for k, (lx, ly, lz) in enumerate(looks):
            y = oracle_ge2.harmonics((cosine * lx - sine * lz, ly, sine * lx + cosine * lz),
                                     max_out)
            for p, (lam, mu) in enumerate(zip(pair_lambda, pair_mu)):
                result[k, point] += alpha[lam] * y[lam * lam + lam + mu] * response[p]
    return result


def near_field_ge2(kernel, elements, bundle, time_origin_ns, *, radius_m=None, panel_m=0.05,
                   gauss_order=6):
    """``(oms, charge (N,), spectrum (W, N))`` of the clamped near portions, the
    spectrum relative to ``time_origin_ns`` (``near_field.py:91-173``)."""
    low = float(np.exp(bundle["log_r"][0]))
    high = float(np.exp(bundle["log_r"][-1]))
    radius = low if radius_m is None else float(radius_m)
    omega = np.asarray(bundle["omega"], float)
    detector = kernel.detector
    positions = np.asarray(detector.positions_m, float)
    n = len(positions)
    origins = np.asarray(time_origin_ns, float)
    wave = kernel.wavelength
    wavelengths = np.asarray(wave.wavelength_nm, float)
    sample = kernel.medium.sample(wavelengths)
    efficiency = np.asarray(detector.spectral_weight(wavelengths), float)
    alpha = np.asarray(kernel.acceptance()["alpha"], float)
    area = np.broadcast_to(np.asarray(detector.effective_area_m2, float), (n,))
    nodes, weights = np.polynomial.legendre.leggauss(int(gauss_order))
    charge = np.zeros(n)
    spectrum = np.zeros((len(omega), n), complex)
    portions = near_portions(positions, elements, radius)
    for om, parts in portions.items():
        position = positions[om]
        normal = np.asarray(detector.orientations[om], float)
        for index, s0, s1 in parts:
            axis = np.asarray(elements.direction[index], float)
            delta = position - np.asarray(elements.start_m[index], float)
            z = float(delta @ axis)
            radial = delta - z * axis
            rho = float(np.linalg.norm(radial))
            if rho < 1e-9:
                trial = np.array([1., 0., 0.]) if abs(axis[0]) < .9 else np.array([0., 1., 0.])
                radial = trial - (trial @ axis) * axis
                rho = 1e-9
            ex = radial / np.linalg.norm(radial)
            ey = np.cross(axis, ex)
            look = -np.array([normal @ ex, normal @ ey, normal @ axis])
            harmonics = cone_harmonics(elements, index, wavelengths, sample["phase_index"],
                                       int(bundle["degree"]))
            if not np.any(harmonics):
                continue
            count = max(1, int(np.ceil((s1 - s0) / panel_m)))
            edges = np.linspace(z - s1, z - s0, count + 1)
            half = np.diff(edges) / 2
            u = (edges[:-1, None] + half[:, None] * (nodes + 1)).ravel()
            w = (half[:, None] * weights).ravel()
            band_weights = (np.asarray(wave.weight_nm, float) * efficiency * float(area[om]))
            values = cone_many_looks(rho, u, look, harmonics, band_weights, alpha,
                                     bundle["pair_lambda"], bundle["pair_mu"], bundle["log_r"],
                                     bundle["coefficients"], bundle["absorption"],
                                     bundle["speed"], omega, bundle["phase"] == "flight",
                                     low, high, 0.0)[0]
            length = float(elements.length_m[index])
            clock = float(elements.end_ns[index] - elements.start_ns[index]) / length
            density = float(elements.coefficient0[index]) / length
            closest = float(elements.start_ns[index]) + z * clock - origins[om]
            part = density * np.exp(1j * omega * closest) * (
                (w[:, None] * values * np.exp(-1j * u[:, None] * clock * omega)).sum(0))
            spectrum[:, om] += part
            charge[om] += part[0].real
    return np.array(sorted(portions), int), charge, spectrum
