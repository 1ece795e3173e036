"""Oracle of stage (a): order-0 direct light to the first glass entry (FP64).

Python side of ``oracle/order0.c``: candidate pairs (``interval_candidates``,
``_ballistic_interval_numba.py:450-519``) and the converged first-entry
charge and arrival bins per OM (``PreparedBallistic.charge``/``.bins``,
``_surface_ballistic.py:91-112``).

Parity pins (tests/test_order0_oracle.py):
  * the axial constant-phase case of ``tests/test_ballistic_first_entry.py:
    27-47`` (closed form in the azimuth, every azimuth hits), charge and bins;
  * the dispersive tilted receiver of ``tests/test_ballistic_first_entry.py:
    77-123`` against an independent brute-force (wavelength, source, azimuth)
    integral written in the test (reference tolerance 2e-6);
  * the two independent evaluators (cap quadrature and the direct-ray
    restatement) agree to 1e-8 on the demonstration event;
  * splitting every element in two leaves the charge unchanged; bins sum to
    the charge when the window holds every arrival.
The reference's omission of pairs whose summed upper bound is below 1.25e-11
pe per OM (``_surface_ballistic.py:60-77``) is not applied: this oracle keeps
every candidate pair.
"""
import ctypes

import numpy as np

from ._clib import arr, lib

_i64p = ctypes.POINTER(ctypes.c_int64)


def _elements_arrays(elements):
    return dict(start=np.ascontiguousarray(elements.start_m, float),
                direction=np.ascontiguousarray(elements.direction, float),
                length=np.ascontiguousarray(elements.length_m, float),
                c0=np.ascontiguousarray(elements.coefficient0, float),
                c2=np.ascontiguousarray(elements.coefficient2, float),
                beta=np.ascontiguousarray(elements.beta, float),
                t0=np.ascontiguousarray(elements.start_ns, float),
                t1=np.ascontiguousarray(elements.end_ns, float))


def candidates(positions, radii, elements, optics):
    """Receiver-major candidate pairs (om, element) of the first-entry route."""
    e = _elements_arrays(elements)
    pos, p_pos = arr(positions, np.float64)
    rad, p_rad = arr(np.broadcast_to(radii, (len(pos),)), np.float64)
    phase = np.asarray(optics["phase_nodes"], float)
    capacity = 1 << 16
    f = lib().oracle_order0_candidates
    f.restype = ctypes.c_int64
    while True:
        om = np.empty(capacity, np.int64)
        el = np.empty(capacity, np.int64)
        count = f(p_pos, p_rad, ctypes.c_int(len(pos)), e["start"].ctypes.data_as(ctypes.c_void_p),
                  e["direction"].ctypes.data_as(ctypes.c_void_p),
                  e["length"].ctypes.data_as(ctypes.c_void_p), e["c0"].ctypes.data_as(ctypes.c_void_p),
                  e["beta"].ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(e["length"])),
                  ctypes.c_double(phase.min()), ctypes.c_double(phase.max()),
                  om.ctypes.data_as(ctypes.c_void_p), el.ctypes.data_as(ctypes.c_void_p),
                  ctypes.c_int64(capacity))
        if count < 0:
            raise ValueError('emitting segment intersects the glass boundary')
        if count <= capacity:
            return om[:count].copy(), el[:count].copy()
        capacity = int(count)


def first_entry(positions, looks, areas, radii, elements, optics, alpha, *, origins=None,
                edges=None, rel_tol=1e-11, bin_tol=5e-5, force_rays=False, pairs=None,
                return_uncertainty=False):
    """Charge (OM,) and bins (OM, bin) or None; looks are ``-orientations``.

    Every pair's charge converges to ``rel_tol``; its bins to ``bin_tol`` of
    its charge or, at the highest order, keep their last L1 change, summed per
    OM into the bin uncertainty returned with ``return_uncertainty``.
    """
    positions = np.ascontiguousarray(positions, float)
    nom = len(positions)
    radii = np.ascontiguousarray(np.broadcast_to(radii, (nom,)), float)
    areas = np.ascontiguousarray(np.broadcast_to(areas, (nom,)), float)
    looks = np.ascontiguousarray(looks, float)
    if pairs is None:
        pairs = candidates(positions, radii, elements, optics)
    po, pe = (np.ascontiguousarray(p, np.int64) for p in pairs)
    e = _elements_arrays(elements)
    want_bins = edges is not None
    edges_a = np.ascontiguousarray(edges if want_bins else [0., 1.], float)
    orig = np.ascontiguousarray(origins if origins is not None else np.zeros(nom), float)
    nbins = len(edges_a) - 1
    charge = np.zeros(nom)
    bins = np.zeros((nom, nbins)) if want_bins else None
    change = np.zeros(nom)
    eff = np.ascontiguousarray(optics["efficiency_coefficients"], float)
    al = np.ascontiguousarray(alpha, float)
    keep = [positions, looks, areas, radii, po, pe, edges_a, orig, eff, al]
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    opt = {k: np.ascontiguousarray(optics[k], float) for k in
           ("panel_edges", "phase_nodes", "group_nodes", "extinction_nodes")}
    f = lib().oracle_order0
    f.restype = ctypes.c_int
    failed = f(ptr(po), ptr(pe), ctypes.c_int64(len(po)), ptr(positions), ptr(looks), ptr(areas),
               ptr(radii), ctypes.c_int(nom), ptr(e["start"]), ptr(e["direction"]),
               ptr(e["length"]), ptr(e["c0"]), ptr(e["c2"]), ptr(e["beta"]), ptr(e["t0"]),
               ptr(e["t1"]), ctypes.c_int(len(opt["panel_edges"]) - 1), ptr(opt["panel_edges"]),
               ptr(opt["phase_nodes"]), ptr(opt["group_nodes"]), ptr(opt["extinction_nodes"]),
               ptr(eff), ctypes.c_int(eff.shape[1]), ptr(al), ctypes.c_int(len(al)),
               ptr(orig) if want_bins else None, ptr(edges_a) if want_bins else None,
               ctypes.c_int(nbins), ctypes.c_double(rel_tol), ctypes.c_double(bin_tol),
               ctypes.c_int(int(force_rays)),
               ptr(charge), ptr(bins) if want_bins else None, ptr(change))
    del keep
    if failed:
        raise ValueError("order-0 oracle quadrature did not converge for some pair")
    if return_uncertainty:
        return charge, bins, change
    return charge, bins
