"""Oracle of the shower route (``transport_shower``) in FP64 — TEST INFRASTRUCTURE ONLY.

Restates ``shower/api.py:223-539`` with the oracle's stages: node-band sources
compiled by ``oracle.compile`` (``shower/ge2.py:21-67``), the fold re-fit of
``oracle.fold``, windows and far-tail screen (``shower_basis.py:388-439``), the
collapsed predictor (``shower/ge2.py:70-96``), orders 0 and 1 of
``transport_prompt`` (``prompt.py:500-947``) from ``oracle.order0`` and
``oracle.order1``, the exact zero-frequency charges and fallback, the full
apply of ``oracle.ge2`` and the spline readout of ``oracle.readout``; the
host preparations (frame, optical panels, node bands, segment bounds, delay
groups) come from ``oracle.host``. Delayed light (``shower/api.py:468-477``),
local time origins (``:356-375``), the predictor fallback (``:405-416``) and the
module selection of the prompt (``prompt.py:596-626``) are restated too.

Pins: every stage is pinned in its own tests; this module is the composition
that the GPU route must reproduce (tests/test_route_gpu.py).
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np

from . import compile as oracle_compile
from . import fold as oracle_fold
from . import host as oracle_host
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


def _selection(positions, elements, area, kernel, tables, threshold, max_distance, min_distance_m,
               peak_acceptance):
    """OMs inside ``max_distance`` of the source and the guard on the others
    (``prompt.py:596-626``)."""
    distance = oracle_host.min_distance(positions, elements)
    inside = distance <= max_distance
    density = (elements.coefficient0[:, None] / tables["wavelength_nm"][None, :] ** 2
               + elements.coefficient2[:, None]
               / (tables["wavelength_nm"][None, :] ** 2 * tables["phase_index"][None, :] ** 2))
    emitted = float(np.sum(tables["weight_nm"][None, :] * np.maximum(density, 0.0)))
    far = np.maximum(distance, 1e-9)
    upper = (10 * max(emitted, 0.0) * area * peak_acceptance * float(np.max(tables["efficiency"]))
             * np.exp(-float(np.min(tables["mua"])) * far) / (4 * np.pi * far ** 2))
    if np.any(~inside & (upper >= threshold)):
        raise ValueError("OMs beyond max_distance_m have a conservative estimate above "
                         "threshold; increase max_distance_m")
    if np.any(inside & (distance < min_distance_m)):
        raise ValueError("OMs are closer than min_distance_m to the source")
    return inside


def transport_shower(kernel, elements, bundle, bands=None, *, config, subdivide=1,
                     time_origin_ns=None, selection_extra_charge=None, predictor_scale=1.0):
    """Route response arrays of one event: dict with ``charge`` (N, 3), ``bins``
    (N, nb, 3), ``spectrum`` (W, N), ``origins``, ``active``, ``level``, ``fallback``.

    The node bands are the oracle's own choice, or ``bands`` (a tuple, or any object
    with a ``bands`` attribute) where the best subsets tie to rounding (see
    ``oracle.host.node_bands``); the interpolation is computed here. ``predictor_scale``
    scales the collapsed predictor's channels (1: the route; a tiny value forces
    the exact-everywhere fallback, as ``tests/test_shower_route.py:161-178`` does
    by replacing ``collapse``)."""
    basis = config.basis
    omega = np.asarray(kernel.config.omega_per_ns, float)
    wavelengths = np.asarray(kernel.wavelength.wavelength_nm, float)
    sample = kernel.medium.sample(wavelengths)
    max_phase = float(np.max(sample["phase_index"]))
    elements = elements.subset(elements.beta * max_phase > 1)
    efficiency = np.asarray(kernel.detector.spectral_weight(wavelengths), float)
    detector = kernel.detector
    positions = np.asarray(detector.positions_m, float)
    n = len(positions)
    orientation = np.asarray(detector.orientations, float)
    looks = -orientation / np.linalg.norm(orientation, axis=1)[:, None]
    area = np.broadcast_to(np.asarray(detector.effective_area_m2, float), (n,)).copy()
    radii = np.broadcast_to(detector.receiver_surface.radii(n), (n,)).copy()
    alpha = np.asarray(kernel.acceptance()["alpha"], float)
    low, high = kernel.config.radial_range_m
    threshold = float(kernel.config.threshold_pe)
    edges = np.asarray(kernel.config.relative_time_edges_ns, float)
    degree, azimuthal = int(kernel.config.source_degree), int(kernel.config.azimuthal_degree)
    extra = (np.zeros(n) if selection_extra_charge is None
             else np.maximum(np.asarray(selection_extra_charge, float), 0.0))
    fold = oracle_host.node_bands(wavelengths, kernel.wavelength.weight_nm, efficiency,
                                  sample["phase_index"], degree, int(config.band_rank),
                                  bands=getattr(bands, "bands", bands))
    folded = oracle_fold.fold_bundle(bundle, fold, subdivide=subdivide)
    folded["coefficients"] = folded["coefficients"].astype(np.complex64)
    zero = dict(folded, coefficients=np.ascontiguousarray(folded["coefficients"][..., :1]),
                omega=folded["omega"][:1])
    group_max = float(np.max(sample["group_index"]))
    prompt_mask, late_groups, beyond = oracle_host.delay_groups(
        elements, prompt_delay_ns=float(config.prompt_delay_ns),
        cell_ns=float(config.delayed_cell_ns), window_end_ns=float(edges[-1]),
        group_index_max=group_max)

    def compiled_part(mask, frequencies):
        part = elements if np.all(mask) else elements.subset(mask)
        frame = oracle_host.axis_frame(part.field(0, phase_index=max_phase))
        return _node_sources(part, wavelengths, fold, frequencies, frame, degree, azimuthal,
                             basis)

    sources = compiled_part(prompt_mask, omega)
    late_zero = compiled_part(~prompt_mask, omega[:1]) if np.any(~prompt_mask) else None
    late_sources = [(shift, compiled_part(mask, omega)) for mask, shift in late_groups]
    # radial windows of every point set (prompt and delayed)
    used = np.zeros(n, np.int64)
    near = np.zeros(n, bool)
    for group in [sources] + ([late_zero] if late_zero is not None else []):
        for item in group:
            radius = np.linalg.norm(item.points_m()[None] - positions[:, None], axis=2)
            used = np.maximum(used, np.count_nonzero((radius >= low) & (radius <= high), axis=1))
            near |= np.any(radius < low, axis=1)
    reach = np.flatnonzero(used > 0)
    predicted = np.zeros(n)
    late_charge = np.zeros(n)
    collapsed = [replace(item, channels=item.channels * predictor_scale)
                 for item in _collapse(sources, float(config.predictor_cell_m))]
    if len(reach):
        predicted[reach] = oracle_ge2.apply_folded(zero, collapsed, positions[reach], looks[reach],
                                                   alpha, area[reach], (low, high))[0].real
        if late_zero is not None:
            late_charge[reach] = oracle_ge2.apply_folded(
                zero, late_zero, positions[reach], looks[reach], alpha, area[reach],
                (low, high))[0].real
    # time origins (shower/api.py:356-375)
    fastest = float(np.min(sample["group_index"]))
    midpoints = elements.start_m + 0.5 * elements.length_m[:, None] * elements.direction
    extent = float(np.linalg.norm(midpoints.max(axis=0) - midpoints.min(axis=0)))
    span = float(np.max(elements.end_ns) - np.min(elements.start_ns))
    local_origin = extent > high or span > 0.5 * float(edges[-1] - edges[0])
    if time_origin_ns is not None:
        origins = np.asarray(time_origin_ns, float)
        local_origin = True
    elif local_origin:
        origins = oracle_host.earliest_arrival(positions, elements, fastest)
    else:
        first = int(np.argmin(elements.start_ns))
        origins = (float(elements.start_ns[first])
                   + np.linalg.norm(positions - elements.start_m[first], axis=1) / (C / fastest))
    # prompt orders 0 and 1 with the predicted order >=2 in the selection
    safety = float(config.predictor_safety)
    selection = safety * np.abs(predicted) + np.abs(late_charge) + extra
    tables = oracle_order1.spectral_tables(kernel.medium, detector, kernel.wavelength)
    peak = float(np.max(np.abs(detector.angular_acceptance(np.linspace(-1, 1, 257)))))
    prompt_settings = config.prompt
    max_distance = (float(high) if getattr(prompt_settings, "max_distance_m", None) is None
                    else float(prompt_settings.max_distance_m))
    inside = _selection(positions, elements, area, kernel, tables, threshold, max_distance,
                        float(getattr(prompt_settings, "min_distance_m", 0.1)), peak)
    work = np.flatnonzero(inside)
    optics = oracle_host.prepare_optics(kernel.medium, detector, kernel.config.wavelength_range_nm)
    q0 = np.zeros(n)
    if len(work):
        q0[work], _ = oracle_order0.first_entry(positions[work], looks[work], area[work],
                                                radii[work], elements, optics, alpha,
                                                rel_tol=1e-9)
    settings = oracle_order1.PromptSettings()
    g = float(kernel.medium.g)
    q1, bins1, level1 = oracle_order1.spectral_order1(
        elements, positions, looks, area, origins, work, tables, alpha, g, edges, threshold, q0,
        selection, settings)
    charge = np.zeros((n, 3))
    charge[:, 0], charge[:, 1] = q0, q1
    prompt_charge = charge[:, :2].sum(axis=1)
    prompt_active = inside & (prompt_charge + selection >= threshold)
    # exact zero frequency where the decision needs it (shower/api.py:388-427)
    computable = (used > 0) | (prompt_charge > 0)
    candidate = (used > 0) & (prompt_charge + safety * np.abs(predicted) + late_charge + extra
                              >= threshold)
    zero_sources = [replace(s, channels=s.channels[:, :, :1]) for s in sources]

    def exact(mask):
        ids = np.flatnonzero(mask)
        values = np.zeros(n)
        if len(ids):
            values[ids] = oracle_ge2.apply_folded(zero, zero_sources, positions[ids], looks[ids],
                                                  alpha, area[ids], (low, high))[0].real
        return values
    exact_charge = exact(candidate)
    fallback = bool(np.any(np.abs(exact_charge[candidate])
                           > safety * np.abs(predicted[candidate])))
    if fallback:
        candidate = used > 0
        exact_charge = exact(candidate)
    level = np.where(used > 0, 1, 0).astype(np.int8)
    charge[:, 2] = predicted + late_charge
    charge[candidate, 2] = exact_charge[candidate] + late_charge[candidate]
    level[candidate] = 2
    active = computable & (charge.sum(axis=1) + extra >= threshold)
    if np.any(active & (level == 1)):
        raise RuntimeError("an OM was activated by a predicted >=2 charge")
    # full frequencies at the active OMs, then the readout (shower/api.py:429-478)
    spectrum = np.zeros((len(omega), n), complex)
    full = np.flatnonzero(active & (used > 0))
    late_spectra = []
    if len(full):
        spectrum[:, full] = oracle_ge2.apply_folded(folded, sources, positions[full], looks[full],
                                                    alpha, area[full], (low, high))
    for shift, group in late_sources:
        values = np.zeros((len(omega), n), complex)
        if len(full):
            values[:, full] = oracle_ge2.apply_folded(folded, group, positions[full],
                                                      looks[full], alpha, area[full], (low, high))
        late_spectra.append((shift, values))
    spectrum[0] = charge[:, 2] - late_charge
    components = np.zeros((n, len(edges) - 1, 3))
    ids = np.flatnonzero(active)
    if len(ids):
        if not np.all(prompt_active[ids]):
            # bins of these OMs were not requested from the prompt (fallback)
            _, bins1, _ = oracle_order1.spectral_order1(
                elements, positions, looks, area, origins, work, tables, alpha, g, edges,
                threshold, q0, None, replace(settings, screen=False))
        _, bins0 = oracle_order0.first_entry(positions[ids], looks[ids], area[ids], radii[ids],
                                             elements, optics, alpha, origins=origins[ids],
                                             edges=edges, rel_tol=1e-9)
        components[ids, :, 0] = bins0
        components[ids, :, 1] = bins1[ids]
        readout_omega = kernel.config.readout_omega_per_ns
        components[ids, :, 2] = oracle_readout.route_readout(
            omega, spectrum[:, ids], origins[ids], readout_omega, edges)
        for shift, values in late_spectra:
            # a delayed group read on its own clock: origin + shift, edges - shift
            components[ids, :, 2] += oracle_readout.route_readout(
                omega, values[:, ids], origins[ids] + shift, readout_omega, edges - shift)
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
                order1_level=level1, near_oms=near_oms, local_origin=local_origin,
                groups_in_window=len(late_sources), late_charge=late_charge,
                fold_bands=tuple(fold.bands), inside=inside)
