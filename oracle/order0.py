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
                edges=None, rel_tol=1e-11, abs_tol=1e-16, bin_tol=5e-5, force_rays=False,
                pairs=None, return_uncertainty=False):
    """Charge (OM,) and bins (OM, bin) or None; looks are ``-orientations``.

    Every pair's charge converges to ``rel_tol`` of itself plus ``abs_tol`` pe
    (pairs of negligible light stop early); its bins to ``bin_tol`` of its
    charge (plus ``abs_tol``) or, at the highest order, keep their last L1
    change, summed per OM into the bin uncertainty returned with
    ``return_uncertainty``.
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
               ctypes.c_int(nbins), ctypes.c_double(rel_tol), ctypes.c_double(abs_tol),
               ctypes.c_double(bin_tol),
               ctypes.c_int(int(force_rays)),
               ptr(charge), ptr(bins) if want_bins else None, ptr(change))
    del keep
    if failed:
        raise ValueError("order-0 oracle quadrature did not converge for some pair")
    if return_uncertainty:
        return charge, bins, change
    return charge, bins


# ---------------------------------------------------------------------------
# The leading finite-cone model (``ballistic_method="interval"``).
#
# ``_ballistic_interval_leading`` (ballistic_interval.py:163-265) with the
# pair integral ``_integral`` (_ballistic_interval_numba.py:189-223), its
# phase-index support ``_support`` (:152-186) and window kinks
# ``_insert_cuts`` (:293-306), as ``_pair_contribution`` (:309-385) applies
# them for charge: the photons leave the element from a window of half-width
# R/sin(theta) around the emission point of the ray through the OM centre,
# weighted by the semicircle law; attenuation, acceptance and distance are
# those of the centre ray. Each OM's Gauss order doubles from 8 until its
# charge changes by at most 1e-10 + rel_tol |q| (ballistic_interval.py:
# 236-251). Not a GPU path: the shower route's default order 0 is the
# first-entry evaluator above. It pins the ingredients the two share
# (elements, optical panels, acceptance, window law) against the reference's
# published demonstration total, which this model produced.

def _acceptance_legendre(alpha, x):
    """``_acceptance`` (_ballistic_interval_numba.py:21-38): sum (2l+1) alpha_l P_l(x) / 4 pi."""
    alpha = np.asarray(alpha, float)
    result = alpha[0] / (4 * np.pi) + (3 * alpha[1] * x / (4 * np.pi) if len(alpha) > 1 else 0.)
    previous, current = np.ones_like(x), x
    for ell in range(2, len(alpha)):
        value = ((2 * ell - 1) * x * current - (ell - 1) * previous) / ell
        result = result + (2 * ell + 1) * alpha[ell] * value / (4 * np.pi)
        previous, current = current, value
    return np.maximum(0., result)


def _semicircle_cdf(x):
    x = np.clip(x, -1., 1.)
    return .5 + (x * np.sqrt(1. - x * x) + np.arcsin(x)) / np.pi


def _support_arrays(z, rho, length, beta, radius):
    """``_support`` for a sphere (radius > 0), vectorised over pairs."""
    quarter = .5 * np.pi
    psi0, psi1 = np.arctan2(rho, z), np.arctan2(rho, z - length)
    a0 = np.arcsin(np.minimum(1., radius / np.hypot(z, rho)))
    a1 = np.arcsin(np.minimum(1., radius / np.hypot(z - length, rho)))
    low, high = psi0 - a0, psi1 + a1
    ok = low < quarter
    floor = (1. + 1e-9) / beta
    with np.errstate(divide="ignore", invalid="ignore"):
        nmin = np.where(low <= 0., floor, np.maximum(floor, 1. / (beta * np.cos(low))))
        nmax = np.where(high >= quarter, np.inf, 1. / (beta * np.cos(high)))
        inner0, inner1 = psi0 + a0, psi1 - a1
        cut0 = np.where(inner0 >= quarter, np.inf, 1. / (beta * np.cos(inner0)))
        cut1 = np.where((inner1 <= 0.) | (inner1 >= quarter), np.inf, 1. / (beta * np.cos(inner1)))
    return ok, nmin, nmax, cut0, cut1


def leading_interval_charge(positions, looks, areas, radii, elements, optics, alpha, *,
                            rel_tol=2e-6, abs_tol=1e-10, order=8, max_order=256, pairs=None):
    """Charge per OM of the leading finite-cone model; looks are ``-orientations``."""
    positions = np.ascontiguousarray(positions, float)
    nom = len(positions)
    radii = np.broadcast_to(np.asarray(radii, float), (nom,))
    areas = np.broadcast_to(np.asarray(areas, float), (nom,))
    looks = np.asarray(looks, float)
    if pairs is None:
        pairs = candidates(positions, radii, elements, optics)
    po, pe = (np.asarray(p, np.int64) for p in pairs)
    u = np.asarray(elements.direction, float)[pe]
    w = positions[po] - np.asarray(elements.start_m, float)[pe]
    z = np.einsum("ij,ij->i", w, u)
    transverse = w - z[:, None] * u
    rho = np.linalg.norm(transverse, axis=1)
    if np.any(rho == 0.):
        raise ValueError("axial pairs need the direct-ray evaluator")
    length = np.asarray(elements.length_m, float)[pe]
    beta = np.asarray(elements.beta, float)[pe]
    c0 = np.asarray(elements.coefficient0, float)[pe]
    c2 = np.asarray(elements.coefficient2, float)[pe]
    radius = radii[po]
    area = areas[po]
    transverse_look = np.einsum("ij,ij->i", transverse, looks[po]) / rho
    axial_look = np.einsum("ij,ij->i", u, looks[po])
    ok, nmin, nmax, cut0, cut1 = _support_arrays(z, rho, length, beta, radius)
    edges = np.asarray(optics["panel_edges"], float)
    phase = np.asarray(optics["phase_nodes"], float)
    ext = np.asarray(optics["extinction_nodes"], float)
    eff = np.asarray(optics["efficiency_coefficients"], float)
    # sub-intervals (pair, a, b, panel): panel support clipped to [nmin, nmax),
    # split at the window kinks cut0, cut1
    rows, lows, highs, panels = [], [], [], []
    for k in range(len(edges) - 1):
        left, width = edges[k], edges[k + 1] - edges[k]
        n0, dn = phase[k], phase[k + 1] - phase[k]
        if dn == 0.:
            raise ValueError("constant-phase panels need the direct-ray evaluator")
        q1, q2 = (nmin - n0) / dn, (nmax - n0) / dn
        a = left + width * np.clip(np.minimum(q1, q2), 0., 1.)
        b = left + width * np.clip(np.maximum(q1, q2), 0., 1.)
        cuts = [a, b]
        for value in (cut0, cut1):
            lam = np.where(np.isfinite(value), left + width * (value - n0) / dn, np.nan)
            cuts.append(np.where((lam > a) & (lam < b), lam, np.nan))
        cuts = np.sort(np.column_stack(cuts), axis=1)   # NaN sort last
        for j in range(3):
            lo, hi = cuts[:, j], cuts[:, j + 1]
            use = ok & (b > a) & np.isfinite(lo) & np.isfinite(hi) & (hi > lo)
            idx = np.flatnonzero(use)
            rows.append(idx)
            lows.append(lo[idx])
            highs.append(hi[idx])
            panels.append(np.full(len(idx), k))
    rows, lows, highs, panels = map(np.concatenate, (rows, lows, highs, panels))

    def evaluate(ng, mask):
        gx, gw = np.polynomial.legendre.leggauss(ng)
        use = mask[po[rows]]
        r, lo, hi, k = rows[use], lows[use], highs[use], panels[use]
        left, width = edges[k], edges[k + 1] - edges[k]
        half, middle = (hi - lo) * .5, (hi + lo) * .5
        lam = middle[:, None] + half[:, None] * gx[None, :]
        q = (lam - left[:, None]) / width[:, None]
        n = phase[k][:, None] + (phase[k + 1] - phase[k])[:, None] * q
        mu = ext[k][:, None] + (ext[k + 1] - ext[k])[:, None] * q
        bn = beta[r][:, None] * n
        valid = bn > 1.
        v = np.sqrt(np.maximum((bn - 1.) * (bn + 1.), 1e-300))
        sine, cosine = v / bn, 1. / bn
        centre = z[r][:, None] - rho[r][:, None] / v
        spread = radius[r][:, None] / sine
        window = (_semicircle_cdf((length[r][:, None] - centre) / spread)
                  - _semicircle_cdf(-centre / spread))
        distance = rho[r][:, None] / sine
        angle = np.clip(transverse_look[r][:, None] * sine + axial_look[r][:, None] * cosine,
                        -1., 1.)
        e = eff[k]
        response = ((e[:, 3:4] * q + e[:, 2:3]) * q + e[:, 1:2]) * q + e[:, 0:1]
        photons = c0[r][:, None] + c2[r][:, None] / (n * n)
        value = (gw[None, :] * np.maximum(0., response) * np.maximum(0., photons) / (lam * lam)
                 * np.exp(-mu * distance) * _acceptance_legendre(alpha, angle)
                 / (2 * np.pi * rho[r][:, None] * sine) * window)
        value = np.where(valid & (window > 0.), value, 0.)
        piece = value.sum(axis=1) * half * area[r] / length[r]
        charge = np.zeros(nom)
        np.add.at(charge, po[r], piece)
        return charge

    pending = np.ones(nom, bool)
    charge = evaluate(order, pending)
    while pending.any():
        if order >= max_order:
            raise ValueError(f"leading interval quadrature failed for {pending.sum()} OMs")
        order = min(max_order, 2 * order)
        q = evaluate(order, pending)
        bad = pending & (np.abs(q - charge) > abs_tol + rel_tol * np.abs(q))
        charge[pending] = q[pending]
        pending = bad
    return charge
