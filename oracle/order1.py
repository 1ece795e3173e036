"""Oracle of order 1 (single scattering) of the shower route, in FP64.

``compress`` restates ``prompt.py:1317-1402`` (+ ``pack_pencils``,
``_prompt_screen.py:44-85``, for the screen's parent/child pencils);
``grouped`` calls the C restatement of ``grouped_pencils_order1`` /
``screen_order1`` (``oracle/order1.c``); ``order1_task`` restates the
grouped branch of ``_order1`` (``prompt.py:1020-1258``) and
``spectral_order1`` the spectral-cone order-1 block of ``transport_prompt``
(``prompt.py:670-851``) with ``PromptConfig.fuse_bands`` units
(``muon/secondary_first.py:184-204``).

Parity pins (tests/test_order1_oracle.py):
  * pencils conserve both photon fields and the first moments of the rings
    (exact sums over ring nodes, independent of the compression);
  * the screen hierarchy conserves each parent's emission in its children;
  * the grouped charge of an isolated point emitter against a direct
    numerical single-scattering integral (Henyey-Greenstein phase, exponential
    attenuation, polynomial acceptance) evaluated independently in the test;
  * bins sum to the charge inside the window, and a rigid time shift of the
    source and origins leaves bins and charge unchanged.
The grouping discretisation itself (cells, pixels, merge widths) has no
independent reference value: the grouped charge is pinned against the direct
integral only to the accuracy of that discretisation (0.2 % off the cone), so
the exact grouped numbers are pinned, beyond that level, by the reference's
published demonstration total (order 1 of the README event, 50.68 pe), which
the restatement reproduces to the quoted digits.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from ._clib import arr, lib

C_M_PER_NS = 0.299792458
NMOM = 23
_GL48 = np.polynomial.legendre.leggauss(48)
GAUSS_TAU = 0.25 * math.pi * (_GL48[0] + 1.0)
GAUSS_W = 0.25 * math.pi * _GL48[1]
TAYLOR = np.asarray([[1.0 / (math.factorial(n) * (n + j + 1.0)) for n in range(24)]
                     for j in range(3)], dtype=np.float64)


@dataclass(frozen=True)
class PromptSettings:
    """The ``PromptConfig`` fields order 1 reads (``prompt.py:67-196``), with the
    shower route's choices (``shower/api.py:36-45``) as defaults."""
    path_step: float = 0.1
    path_step_tail: float = 0.25
    tail_exponent: float = 25.0
    direct_bins: int = 4
    max_path_nodes: int = 2048
    cell_m: float = 0.5
    pixel_face: int = 32
    ring_nodes: int = 96
    smoothing_factor: float = 6.0
    merge_theta_min: float = 2e-3
    merge_theta_ratio: float = 0.08
    merge_tau_ns: float = 0.5
    merge_log_r: float = 0.02
    screen: bool = True
    screen_margin: float = 3.0
    screen_cell_m: float = 2.0
    screen_pixel_face: int = 8
    screen_ring_nodes: int = 24
    screen_refine_factor: int = 4


def acceptance_monomials(alpha):
    """``coef, ipow, jpow, degree`` of ``A(x) = sum alpha_l (2l+1) P_l(x) / (4 pi)``
    in powers of ``cos chi c1`` and ``sin chi c3`` (``prompt.py:380-399``)."""
    alpha = np.asarray(alpha, float)
    series = alpha * (2 * np.arange(len(alpha)) + 1) / (4 * np.pi)
    power = np.polynomial.legendre.leg2poly(series)
    degree = len(power) - 1
    coef, ipow, jpow = [], [], []
    for k in range(degree + 1):
        for i in range(k + 1):
            coef.append(power[k] * math.comb(k, i))
            ipow.append(i)
            jpow.append(k - i)
    return (np.asarray(coef, float), np.asarray(ipow, np.int64), np.asarray(jpow, np.int64),
            int(degree))


def spectral_tables(medium, detector, quadrature):
    """Per-wavelength optics and spectral weights (``prompt.py:412-434``)."""
    wavelengths = np.asarray(quadrature.wavelength_nm, float)
    weights = np.asarray(quadrature.weight_nm, float)
    sample = medium.sample(wavelengths)
    efficiency = np.array([float(detector.spectral_weight(float(w))) for w in wavelengths])
    mus = np.asarray(sample["scattering_per_m"], float)
    mua = np.asarray(sample["absorption_per_m"], float)
    speed = C_M_PER_NS / np.asarray(sample["group_index"], float)
    phase = np.asarray(sample["phase_index"], float)
    return {"wavelength_nm": wavelengths, "weight_nm": weights, "efficiency": efficiency,
            "mus": mus, "mua": mua, "mut": mua + mus, "speed": speed, "phase_index": phase,
            "group_index": np.asarray(sample["group_index"], float),
            "s0w": weights * efficiency / wavelengths ** 2,
            "s2w": weights * efficiency / (wavelengths ** 2 * phase ** 2)}


def _chord_arrays(elements, cone):
    return [np.ascontiguousarray(elements.start_m, float),
            np.ascontiguousarray(elements.direction, float),
            np.ascontiguousarray(elements.length_m, float),
            np.ascontiguousarray(elements.coefficient0, float),
            np.ascontiguousarray(elements.coefficient2, float),
            np.ascontiguousarray(cone, float),
            np.ascontiguousarray(elements.start_ns, float),
            np.ascontiguousarray(elements.end_ns, float)]


def _grid(elements, cell_m):
    start = np.asarray(elements.start_m, float)
    ends = start + np.asarray(elements.length_m, float)[:, None] * np.asarray(elements.direction)
    low = np.minimum(start.min(axis=0), ends.min(axis=0)) - cell_m
    high = np.maximum(start.max(axis=0), ends.max(axis=0)) + cell_m
    dims = np.ceil((high - low) / cell_m).astype(np.int64) + 1
    return low, dims


def _rows(elements, cone, cell_m, pixel_face, ring_nodes):
    """Moment rows per (cell, pixel) key, sorted by key (``compress_subsegments``)."""
    origin, dims = _grid(elements, cell_m)
    if float(np.prod(dims.astype(float))) * 6 * int(pixel_face) ** 2 >= 2.0 ** 62:
        raise ValueError("cell grid too large for 64-bit pencil keys; increase cell_m")
    arrays = _chord_arrays(elements, cone)
    keep = [arr(a, np.float64) for a in arrays]
    o = arr(origin, np.float64)
    d = arr(dims, np.int64)
    L = lib()
    L.oracle_o1_compress.restype = ctypes.c_int64
    count = L.oracle_o1_compress(ctypes.c_int(len(elements)), *[k[1] for k in keep],
                                 ctypes.c_double(cell_m), ctypes.c_int(int(pixel_face)),
                                 ctypes.c_int(int(ring_nodes)), o[1], d[1])
    keys = np.zeros(count, np.int64)
    rows = np.zeros((count, NMOM))
    L.oracle_o1_rows(keys.ctypes.data_as(ctypes.c_void_p), rows.ctypes.data_as(ctypes.c_void_p))
    return keys, rows


_FIELDS = ("position", "direction", "w0", "w2", "time", "major", "var_major", "var_minor",
           "pos_cov")


def finish(rows):
    """Pencils from moment rows (``finish_pencils``)."""
    n = len(rows)
    out = {"position": np.zeros((n, 3)), "direction": np.zeros((n, 3)), "w0": np.zeros(n),
           "w2": np.zeros(n), "time": np.zeros(n), "major": np.zeros((n, 3)),
           "var_major": np.zeros(n), "var_minor": np.zeros(n), "pos_cov": np.zeros((n, 6))}
    rows = np.ascontiguousarray(rows, float)
    L = lib()
    L.oracle_o1_finish.restype = ctypes.c_int64
    kept = L.oracle_o1_finish(ctypes.c_int64(n), rows.ctypes.data_as(ctypes.c_void_p),
                              *[out[k].ctypes.data_as(ctypes.c_void_p) for k in _FIELDS])
    return {k: v[:kept] for k, v in out.items()}


def compress(elements, cone, cell_m, pixel_face, ring_nodes):
    """Fine pencils of ``elements`` with ring cone cosines ``cone`` (``prompt.py:1317``)."""
    _, rows = _rows(elements, cone, cell_m, pixel_face, ring_nodes)
    pencils = finish(rows)
    a_step = 0.5 * cell_m
    n_a = np.maximum(1, np.ceil(np.asarray(elements.length_m) / a_step))
    pencils["ring_nodes_deposited"] = int(n_a.sum() * ring_nodes)
    return pencils


def parent_keys(keys, face, parent_face):
    """Parent indices of nested cube pixels (``_prompt_screen.py:44-53``)."""
    factor = face // parent_face
    cell, pixel = np.divmod(keys, 6 * face * face)
    cube_face, ij = np.divmod(pixel, face * face)
    i, j = np.divmod(ij, face)
    return (cell * (6 * parent_face * parent_face)
            + (cube_face * parent_face + i // factor) * parent_face + j // factor)


def compress_screen(elements, cone, settings):
    """Parent/child pencils of the coarse screen in preorder
    (``compress_screen`` + ``pack_pencils``, ``_prompt_screen.py:56-93``)."""
    face = int(settings.screen_pixel_face) * int(settings.screen_refine_factor)
    parent_face = int(settings.screen_pixel_face)
    keys, rows = _rows(elements, cone, settings.screen_cell_m, face, settings.screen_ring_nodes)
    fine = finish(rows)
    coarse_of = parent_keys(keys, face, parent_face)
    order = np.lexsort((keys, coarse_of))
    coarse_keys, start = np.unique(coarse_of[order], return_index=True)
    coarse_rows = np.add.reduceat(rows[order], start, axis=0) if len(order) else rows[:0]
    coarse = finish(coarse_rows)
    positive = rows[:, 0] > 0
    fine_keys = np.sort(keys[positive])
    coarse_keys = coarse_keys[coarse_rows[:, 0] > 0]
    parents = np.searchsorted(coarse_keys, parent_keys(fine_keys, face, parent_face))
    n = len(coarse_keys)
    if np.any(parents >= n) or not np.array_equal(coarse_keys[parents],
                                                  parent_keys(fine_keys, face, parent_face)):
        raise ValueError("inconsistent coarse angular hierarchy")
    order = np.argsort(np.r_[np.arange(n), parents], kind="stable")
    inverse = np.empty(len(order), np.int64)
    inverse[order] = np.arange(len(order))
    result = {k: np.ascontiguousarray(np.concatenate([coarse[k], fine[k]])[order])
              for k in _FIELDS}
    parent = np.r_[np.full(n, -2), parents].astype(np.int64)[order]
    child = parent >= 0
    parent[child] = inverse[parent[child]]
    next_node = np.arange(1, len(order) + 1, dtype=np.int64)
    roots = np.flatnonzero(parent == -2)
    next_node[roots] = np.r_[roots[1:], len(order)] if len(roots) else []
    result["hierarchy_parent"] = parent
    result["next_node"] = next_node
    return result


def interval_tables(edges, kappa):
    """``e^{kappa t}`` at the edges and ``I_k(kappa width)`` per bin (``interval_table``)."""
    edges = np.ascontiguousarray(edges, float)
    kappa = np.ascontiguousarray(kappa, float)
    nb = len(edges) - 1
    eedge = np.zeros((len(kappa), nb + 1))
    itab = np.zeros((len(kappa), nb, 3))
    lib().oracle_o1_tables(TAYLOR.ctypes.data_as(ctypes.c_void_p),
                           edges.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(nb),
                           kappa.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(kappa)),
                           eedge.ctypes.data_as(ctypes.c_void_p),
                           itab.ctypes.data_as(ctypes.c_void_p))
    return eedge, itab


def grouped(positions, looks, areas, origins, work, pencils, tables, alpha, g, edges, settings,
            merge, *, compute_bins=True):
    """Grouped (or, with hierarchy arrays, screen) order 1 of OMs ``work``:
    ``(charge (N,), bins (N, nb), groups (N,), batches (N,))``."""
    coef, ipow, jpow, degree = acceptance_monomials(alpha)
    n = len(positions)
    nb = len(edges) - 1
    kappa = -tables["mut"] * tables["speed"]
    eedge, itab = interval_tables(edges, kappa)
    tangent = np.tan(GAUSS_TAU)
    weighted = GAUSS_W * (1.0 + tangent * tangent)
    y_g = (1.0 - abs(g)) / (1.0 + abs(g)) if g != 0 else 1.0
    charge = np.zeros(n)
    bins = np.zeros((n, nb))
    groups = np.zeros(n, np.int64)
    batches = np.zeros(n, np.int64)
    keep = []

    def p(a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a.ctypes.data_as(ctypes.c_void_p)
    screen = "hierarchy_parent" in pencils
    work = np.ascontiguousarray(work, np.int64)
    status = lib().oracle_o1_grouped(
        p(positions), p(looks), p(areas), p(origins), p(work, np.int64),
        ctypes.c_int64(len(work)), ctypes.c_int64(len(pencils["w0"])),
        p(pencils["position"]), p(pencils["direction"]), p(pencils["w0"]), p(pencils["w2"]),
        p(pencils["time"]), p(pencils["major"]), p(pencils["var_major"]),
        p(pencils["var_minor"]), p(pencils["pos_cov"]),
        p(pencils["hierarchy_parent"], np.int64) if screen else None,
        p(pencils["next_node"], np.int64) if screen else None,
        p(tangent), p(weighted), ctypes.c_int(len(tangent)), p(coef), p(ipow, np.int64),
        p(jpow, np.int64), ctypes.c_int(len(coef)), ctypes.c_int(degree),
        ctypes.c_double(g), ctypes.c_double(y_g), ctypes.c_double(settings.path_step),
        ctypes.c_double(settings.path_step_tail), ctypes.c_double(settings.tail_exponent),
        ctypes.c_int(len(tables["mus"])), p(tables["mus"]), p(tables["mut"]),
        p(tables["speed"]), p(tables["s0w"]), p(tables["s2w"]), p(edges), ctypes.c_int(nb),
        p(eedge), p(itab), p(TAYLOR), ctypes.c_int(int(settings.direct_bins)),
        ctypes.c_int(int(settings.max_path_nodes)), ctypes.c_double(settings.smoothing_factor),
        ctypes.c_double(merge[0]), ctypes.c_double(merge[1]), ctypes.c_double(merge[2]),
        ctypes.c_double(merge[3]), ctypes.c_double(float(np.max(tables["speed"]))),
        ctypes.c_int(1 if compute_bins else 0), charge.ctypes.data_as(ctypes.c_void_p),
        bins.ctypes.data_as(ctypes.c_void_p), groups.ctypes.data_as(ctypes.c_void_p),
        batches.ctypes.data_as(ctypes.c_void_p))
    if status != 0:
        raise ValueError("acceptance degree too high for the order-1 oracle")
    return charge, bins, groups, batches


def evaluation_units(tables, band_valid, fuse_bands=True):
    """Bands evaluated together (``muon/secondary_first.py:184-204``)."""
    nband = len(tables["wavelength_nm"])
    phase = np.asarray(tables["phase_index"], float)
    single = [dict(bands=[b], valid=band_valid[b], cone=float(phase[b]), fused=False,
                   tables={k: v[b:b + 1] for k, v in tables.items()}) for b in range(nband)]
    if not fuse_bands:
        return single
    everywhere = np.logical_and.reduce(band_valid)
    weight = np.asarray(tables["s0w"], float)
    fused = dict(bands=list(range(nband)), valid=everywhere, fused=True,
                 cone=float(np.sum(weight * phase) / np.sum(weight)), tables=tables)
    for unit in single:
        unit["valid"] = unit["valid"] & ~everywhere
    return [fused] + single


def spectral_order1(elements, positions, looks, areas, origins, work, tables, alpha, g, edges,
                    threshold, order0_charge, selection_extra_charge, settings=None,
                    fused_pixel_face=96, fused_ring_nodes=192):
    """Order 1 of a spectral-cone element source on the grouped path with fused
    bands and the global screen (``prompt.py:670-851``): ``(charge, bins, level)``."""
    settings = settings or PromptSettings()
    n = len(positions)
    nb = len(edges) - 1
    work = np.asarray(work, np.int64)
    charge = np.zeros(n)
    bins = np.zeros((n, nb))
    level = np.zeros(n, np.int8)
    band_valid = [elements.beta * phase > 1 for phase in tables["phase_index"]]
    units = evaluation_units(tables, band_valid, True)
    tasks = []
    for unit in units:
        mask = unit["valid"]
        if not np.any(mask):
            continue
        subset = elements if np.all(mask) else elements.subset(mask)
        cone = 1.0 / (np.asarray(subset.beta, float) * float(unit["cone"]))
        face = fused_pixel_face if unit["fused"] else settings.pixel_face
        ring = fused_ring_nodes if unit["fused"] else settings.ring_nodes
        tasks.append((subset, cone, unit["tables"], face, ring))
    extra = (np.zeros(n) if selection_extra_charge is None
             else np.asarray(selection_extra_charge, float))
    global_screen = selection_extra_charge is not None and settings.screen and threshold > 0
    coarse_merge = (settings.merge_theta_min, 0.3, 2.0, settings.merge_log_r)
    fine_merge = (settings.merge_theta_min, settings.merge_theta_ratio, settings.merge_tau_ns,
                  settings.merge_log_r)
    fine_work = work
    coarse_total = None
    if global_screen:
        for subset, cone, unit_tables, _, _ in tasks:
            coarse = compress_screen(subset, cone, settings)
            q, _, _, _ = grouped(positions, looks, areas, origins, work, coarse, unit_tables,
                                 alpha, g, edges, settings, coarse_merge, compute_bins=False)
            charge += q
        level[work] = 1
        coarse_total = charge.copy()
        fine_work = work[(order0_charge[work] + extra[work]
                          + settings.screen_margin * np.maximum(charge[work], 0.0)) >= threshold]
        charge[fine_work] = 0.0
    pencil_sets = []
    for subset, cone, unit_tables, face, ring in tasks:
        pencils = compress(subset, cone, settings.cell_m, face, ring) if len(fine_work) else None
        pencil_sets.append((pencils, unit_tables))
        if not len(fine_work):
            continue
        q, b, _, _ = grouped(positions, looks, areas, origins, fine_work, pencils, unit_tables,
                             alpha, g, edges, settings, fine_merge)
        charge[fine_work] += q[fine_work]
        bins += b
    level[fine_work] = 2
    if global_screen and len(fine_work):
        ratio = coarse_total[fine_work] / np.maximum(charge[fine_work], 1e-300)
        worst = float(np.min(ratio))
        if worst * settings.screen_margin < 1.0:
            wider = 1.5 / max(worst, 1e-3)
            extra_work = work[(level[work] == 1)
                              & (order0_charge[work] + extra[work]
                                 + wider * np.maximum(coarse_total[work], 0.0) >= threshold)]
            if len(extra_work):
                charge[extra_work] = 0.0
                for pencils, unit_tables in pencil_sets:
                    q, b, _, _ = grouped(positions, looks, areas, origins, extra_work, pencils,
                                         unit_tables, alpha, g, edges, settings, fine_merge)
                    charge[extra_work] += q[extra_work]
                    bins[extra_work] += b[extra_work]
                level[extra_work] = 2
    return charge, bins, level
