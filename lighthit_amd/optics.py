"""Optical panels of the direct-light integral (``ballistic_interval.py:64-161``).

The medium is its piecewise-linear table; the detector spectral efficiency
is represented by verified cubic panels in the normalised panel coordinate
``q`` (ascending powers). Panels start from the medium nodes, the wavelength
range, any supplied knots and a 20 nm grid, and are halved where the cubic
misses the efficiency at seven probe points by more than
``1e-12 + 1e-8 |eta|`` or dips below zero.
"""
from __future__ import annotations

import numpy as np


def prepare_optics(medium, detector, wavelength_range_nm=None, *, spectral_knots_nm=None,
                   relative_tolerance=1e-8, absolute_tolerance=1e-12, max_depth=16):
    low, high = wavelength_range_nm or medium.wavelength_range_nm
    if not medium.wavelength_range_nm[0] <= low < high <= medium.wavelength_range_nm[1]:
        raise ValueError('spectral range lies outside medium table')
    supplied = spectral_knots_nm
    if supplied is None:
        supplied = getattr(detector.spectral_efficiency, 'wavelength_knots_nm', None)
    extra = np.asarray([] if supplied is None else supplied, float)
    if extra.ndim != 1 or not np.isfinite(extra).all():
        raise ValueError('spectral response knots must be a finite vector')
    edges = np.unique(np.r_[low, high, medium.wavelength_nm, extra,
                            np.linspace(low, high, max(2, int(np.ceil((high - low) / 20)) + 1))])
    edges = edges[(edges >= low) & (edges <= high)]
    nodes = np.array([0., 1 / 3, 2 / 3, 1.])
    probes = np.array([1 / 8, 1 / 6, 3 / 8, .5, 5 / 8, 5 / 6, 7 / 8])
    inverse = np.linalg.inv(np.polynomial.polynomial.polyvander(nodes, 3))
    probe_basis = np.polynomial.polynomial.polyvander(probes, 3).T
    for depth in range(max_depth + 1):
        widths = np.diff(edges)
        values = np.asarray(detector.spectral_weight(edges[:-1, None] + widths[:, None] * nodes), float)
        coefficients = values @ inverse.T
        exact = np.asarray(detector.spectral_weight(edges[:-1, None] + widths[:, None] * probes), float)
        predicted = coefficients @ probe_basis
        tolerance = absolute_tolerance + relative_tolerance * np.abs(exact)
        bad = np.max(np.abs(predicted - exact) / tolerance, axis=1) > 1
        for i, coef in enumerate(coefficients):
            roots = np.polynomial.polynomial.polyroots(np.array([coef[1], 2 * coef[2], 3 * coef[3]]))
            internal = [r.real for r in roots if abs(r.imag) < 1e-12 and 0 < r.real < 1]
            minimum = min(np.polynomial.polynomial.polyval(np.r_[0., internal, 1.], coef))
            bad[i] |= minimum < -absolute_tolerance
        if not bad.any():
            break
        if depth == max_depth:
            raise ValueError('spectral-efficiency interpolation did not converge; supply its breakpoints')
        edges = np.sort(np.r_[edges, (edges[:-1] + widths / 2)[bad]])
    sample = medium.sample(edges)
    return dict(panel_edges=np.ascontiguousarray(edges),
                phase_nodes=np.ascontiguousarray(sample['phase_index']),
                group_nodes=np.ascontiguousarray(sample['group_index']),
                extinction_nodes=np.ascontiguousarray(sample['absorption_per_m']
                                                      + sample['scattering_per_m']),
                efficiency_coefficients=np.ascontiguousarray(coefficients),
                medium_nodes_nm=np.asarray(medium.wavelength_nm, float).copy())
