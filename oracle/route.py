"""Oracle of the shower route (``transport_shower``) in FP64 — TEST INFRASTRUCTURE ONLY.

Restates ``shower/api.py:223-539`` with the oracle's stages: node-band sources
compiled by ``oracle.compile`` (``shower/ge2.py:21-67``), the fold re-fit of
``oracle.fold``, windows and far-tail screen (``shower_basis.py:388-439``), the
collapsed predictor (``shower/ge2.py:70-96``), orders 0 and 1 of
``transport_prompt`` (``prompt.py:500-947``) from ``oracle.order0`` and
``oracle.order1``, the exact zero-frequency charges and fallback, the full
apply of ``oracle.ge2`` and the spline readout of ``oracle.readout``. The
node-band choice (``BandFold.of``) is the package's host code, pinned by its
own invariant tests.

Pins: every stage is pinned in its own tests; this module is the composition
that the GPU route must reproduce (tests/test_route_gpu.py).
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np

from . import compile as oracle_compile
from . import fold as oracle_fold
from . import ge2 as oracle_ge2
from . import order0 as oracle_order0
from . import order1 as oracle_order1
from . import readout as oracle_readout

C = 0.299792458


def _node_sources(elements, wavelengths, fold, omega, frame, degree, azimuthal, basis):
    compiled = {}
    for band in fold.bands:
        n = float(fold.phase_index[band])
        mask = elements.beta * n > 1
        if not np.any(mask):
            continue
        part = elements if np.all(mask) else elements.subset(mask)
        first, second = part.field(0, phase_index=n), part.field(2, phase_index=n)
        scale0 = 1.0 / float(wavelengths[band]) ** 2
        scale2 = scale0 / n ** 2
        field = replace(first, photons=scale0 * first.photons - scale2 * second.photons)
        compiled[band] = oracle_compile.compile_axial_slices(
            field, degree, omega, azimuthal_degree=azimuthal, cell_m=basis.cell_m,
            axial_order=basis.axial_order, transverse=basis.transverse,
            stencil_m=basis.stencil_m, stencil_threshold_m=basis.stencil_threshold_m,
            time_step_ns=basis.time_step_ns, frame=frame)
    template = next(iter(compiled.values()))
    out = []
    for band in fold.bands:
        item = compiled.get(band, replace(template, channels=np.zeros_like(template.channels)))
        out.append(replace(item, channels=item.channels * fold.scale[band]))
    return out


def _collapse(sources, cell_m):
    reduced = []
    for item in sources:
        weight = np.abs(item.channels[:, 0, 0])
        keys = np.floor(np.column_stack((item.z_m, item.transverse_m)) / cell_m).astype(np.int64)
        _, inverse = np.unique(keys, axis=0, return_inverse=True)
        inverse = np.asarray(inverse).ravel()
        z, t, ch = [], [], []
        for g in range(int(inverse.max()) + 1):
            chunk = np.flatnonzero(inverse == g)
            total = weight[chunk].sum()
            share = weight[chunk] / total if total > 0 else np.full(len(chunk), 1 / len(chunk))
            z.append(np.sum(share * item.z_m[chunk]))
            t.append(share @ item.transverse_m[chunk])
            ch.append(item.channels[chunk, :, :1].sum(axis=0))
        reduced.append(replace(item, z_m=np.asarray(z), transverse_m=np.asarray(t),
                               channels=np.asarray(ch)))
    return reduced


def _attenuation(kernel):
    wavelengths = np.asarray(kernel.wavelength.wavelength_nm, float)
    weights = (np.asarray(kernel.wavelength.weight_nm, float)
               * np.asarray(kernel.detector.spectral_weight(wavelengths), float) / wavelengths ** 2)
    absorption = np.asarray(kernel.medium.sample(wavelengths)["absorption_per_m"], float)
    lowest = float(absorption.min())
    radius = np.geomspace(1e-3, 1e5, 1025)
    log_factor = (np.log(np.exp(-np.outer(radius, absorption - lowest)) @ weights
                         / weights.sum()) - lowest * radius)
    return lambda d: np.exp(np.interp(d, radius, log_factor))


def window_report(points_m, photons, positions, area, low, high, radius, log_factor,
                  peak_acceptance, efficiency):
    """Basis points below, inside and beyond the radial window per OM and the
    far-tail estimate (``shower_basis.py:408-439``), one OM at a time:
    ``(used, far, near, tail)``."""
    n = len(positions)
    used, far, near = (np.zeros(n, np.int64) for _ in range(3))
    tail = np.zeros(n)
    for d in range(n):
        r = np.sqrt(np.sum((points_m - positions[d]) ** 2, axis=1))
        is_near, is_far = r < low, r > high
        near[d], far[d] = int(is_near.sum()), int(is_far.sum())
        used[d] = len(points_m) - near[d] - far[d]
        rf = r[is_far]
        tail[d] = (10 * area[d] * peak_acceptance * efficiency
                   * np.sum(photons[is_far] * np.exp(np.interp(rf, radius, log_factor))
                            / (4 * np.pi * rf ** 2)))
    return used, far, near, tail


def transport_shower(kernel, elements, bundle, fold, *, config, subdivide=1):
    """Route response arrays of one event: dict with ``charge`` (N, 3), ``bins``
    (N, nb, 3), ``spectrum`` (W, N), ``origins``, ``active``, ``level``."""
    from lighthit_amd.axial import AxisFrame
    from lighthit_amd.optics import prepare_optics
    basis = config.basis
    omega = np.asarray(kernel.config.omega_per_ns, float)
    folded = oracle_fold.fold_bundle(bundle, fold, subdivide=subdivide)
    folded["coefficients"] = folded["coefficients"].astype(np.complex64)
    zero = dict(folded, coefficients=np.ascontiguousarray(folded["coefficients"][..., :1]),
                omega=folded["omega"][:1])
    wavelengths = np.asarray(kernel.wavelength.wavelength_nm, float)
    sample = kernel.medium.sample(wavelengths)
    max_phase = float(np.max(sample["phase_index"]))
    elements = elements.subset(elements.beta * max_phase > 1)
    efficiency = np.asarray(kernel.detector.spectral_weight(wavelengths), float)
    detector = kernel.detector
    positions = np.asarray(detector.positions_m, float)
    looks = detector.looks
    area = np.asarray(detector.effective_area_m2, float)
    alpha = np.asarray(kernel.acceptance()["alpha"], float)
    low, high = kernel.config.radial_range_m
    threshold = float(kernel.config.threshold_pe)
    n = len(positions)
    edges = np.asarray(kernel.config.relative_time_edges_ns, float)
    degree, azimuthal = int(kernel.config.source_degree), int(kernel.config.azimuthal_degree)
    frame = AxisFrame.of(elements.field(0, phase_index=max_phase))
    sources = _node_sources(elements, wavelengths, fold, omega, frame, degree, azimuthal, basis)
    # windows
    attenuation = _attenuation(kernel)
    peak = float(np.max(np.abs(detector.angular_acceptance(np.linspace(-1, 1, 257)))))
    used = np.zeros(n, np.int64)
    near = np.zeros(n, bool)
    for band, item in zip(fold.bands, sources):
        radius = np.linalg.norm(item.points_m()[None] - positions[:, None], axis=2)
        used = np.maximum(used, np.count_nonzero((radius >= low) & (radius <= high), axis=1))
        near |= np.any(radius < low, axis=1)
    reach = np.flatnonzero(used > 0)
    predicted = np.zeros(n)
    collapsed = _collapse(sources, float(config.predictor_cell_m))
    if len(reach):
        predicted[reach] = oracle_ge2.apply_folded(zero, collapsed, positions[reach], looks[reach],
                                                   alpha, area[reach], (low, high))[0].real
    # origins
    fastest = float(np.min(sample["group_index"]))
    first = int(np.argmin(elements.start_ns))
    origins = (float(elements.start_ns[first])
               + np.linalg.norm(positions - elements.start_m[first], axis=1) / (C / fastest))
    # prompt: orders 0 and 1 with the predicted order >=2 in the selection
    safety = float(config.predictor_safety)
    extra = safety * np.abs(predicted)
    tables = oracle_order1.spectral_tables(kernel.medium, detector, kernel.wavelength)
    optics = prepare_optics(kernel.medium, detector)
    q0, _ = oracle_order0.first_entry(positions, looks, area, detector.receiver_surface.radii(n),
                                      elements, optics, alpha, rel_tol=1e-9)
    settings = oracle_order1.PromptSettings()
    q1, bins1, level1 = oracle_order1.spectral_order1(
        elements, positions, looks, area, origins, np.arange(n), tables, alpha,
        float(kernel.medium.g), edges, threshold, q0, extra, settings)
    charge = np.zeros((n, 3))
    charge[:, 0], charge[:, 1] = q0, q1
    prompt_active = charge[:, :2].sum(axis=1) + extra >= threshold
    # exact zero frequency where the decision needs it
    prompt_charge = charge[:, :2].sum(axis=1)
    computable = (used > 0) | (prompt_charge > 0)
    candidate = (used > 0) & (prompt_charge + safety * np.abs(predicted) >= threshold)

    def exact(mask):
        ids = np.flatnonzero(mask)
        values = np.zeros(n)
        if len(ids):
            values[ids] = oracle_ge2.apply_folded(zero, [replace(s, channels=s.channels[:, :, :1])
                                                         for s in sources], positions[ids],
                                                  looks[ids], alpha, area[ids], (low, high))[0].real
        return values
    exact_charge = exact(candidate)
    fallback = bool(np.any(np.abs(exact_charge[candidate])
                           > safety * np.abs(predicted[candidate])))
    if fallback:
        candidate = used > 0
        exact_charge = exact(candidate)
    level = np.where(used > 0, 1, 0).astype(np.int8)
    charge[:, 2] = predicted
    charge[candidate, 2] = exact_charge[candidate]
    level[candidate] = 2
    active = computable & (charge.sum(axis=1) >= threshold)
    spectrum = np.zeros((len(omega), n), complex)
    work = np.flatnonzero(active & (used > 0))
    if len(work):
        spectrum[:, work] = oracle_ge2.apply_folded(folded, sources, positions[work], looks[work],
                                                    alpha, area[work], (low, high))
    spectrum[0] = charge[:, 2]
    components = np.zeros((n, len(edges) - 1, 3))
    ids = np.flatnonzero(active)
    if len(ids):
        _, bins0 = oracle_order0.first_entry(positions[ids], looks[ids], area[ids],
                                             detector.receiver_surface.radii(n)[ids], elements,
                                             optics, alpha, origins=origins[ids], edges=edges,
                                             rel_tol=1e-9)
        components[ids, :, 0] = bins0
        components[ids, :, 1] = bins1[ids]
        components[ids, :, 2] = oracle_readout.route_readout(
            omega, spectrum[:, ids], origins[ids], kernel.config.readout_omega_per_ns, edges)
    near_oms = np.zeros(0, int)
    if np.any(near):
        # the near-field clamp (shower/api.py:536-538, near_field.py:176-217)
        from . import near_field as oracle_near
        near_oms, near_charge, near_spectrum = oracle_near.near_field_ge2(
            kernel, elements, bundle, origins)
        charge[near_oms, 2] += near_charge[near_oms]
        detailed = near_oms[active[near_oms]]
        if len(detailed):
            components[detailed, :, 2] += oracle_readout.route_readout(
                omega, near_spectrum[:, detailed], np.zeros(len(detailed)),
                kernel.config.readout_omega_per_ns, edges)
        spectrum[:, near_oms] += near_spectrum[:, near_oms] * np.exp(
            1j * omega[:, None] * origins[None, near_oms])
    return dict(charge=charge, bins=components, spectrum=spectrum, origins=origins,
                active=active, level=level, prompt_active=prompt_active, fallback=fallback,
                order1_level=level1, near_oms=near_oms)
