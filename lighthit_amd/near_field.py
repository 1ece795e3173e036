"""Order >=2 light of source portions closer to an OM than the tables reach.

The near-field clamp of ``near_field.py:1-217``: the order >=2 tables start at
their radial minimum; element portions inside it are left out of the route's
sums and added back here with the amplitude at the minimum along the same
directions and the flight phase of the true distance
(``cone_point_evaluator(..., near_field=True)``,
``muon/track_ge2_primitive.py:84-162, 612-660``). The portions and their frames
are set up on the host; ``lh_near_field`` evaluates them on the GPU. Every OM
that uses the clamp is reported in ``metadata['near_field']`` and named in a
``RuntimeWarning``.
"""
from __future__ import annotations

import ctypes
import warnings
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .ge2 import _Pack


class _Portions(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("impact", "look", "weight", "harmonics",
                                               "phase_ns", "clock", "density", "node_start",
                                               "u", "w")] + [("count", ctypes.c_int)]


_lib.SIGNATURES["lh_near_field"] = (ctypes.c_int, [
    ctypes.POINTER(_Pack), ctypes.POINTER(_Portions), ctypes.c_void_p, ctypes.c_int,
    ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p])


@dataclass
class NearFieldGe2:
    """Clamped order >=2 of the near source portions, per OM (``near_field.py:30-36``)."""
    oms: np.ndarray                 # OM indices that have near source points
    charge_pe: np.ndarray           # (n,) clamped order >=2 charge
    spectrum_pe: np.ndarray         # (frequency, n) relative to the OM time origins
    metadata: dict = field(default_factory=dict)


def near_portions(positions, elements, radius, candidates=None):
    """{om: [(element, s0, s1), ...]} of the element portions within ``radius`` of
    each OM, in element order (``near_field.py:44-69``). ``candidates`` limits the
    OMs searched (those whose nearest segment lies within ``radius``)."""
    positions = np.asarray(positions, float)
    start = np.asarray(elements.start_m, float)
    direction = np.asarray(elements.direction, float)
    length = np.asarray(elements.length_m, float)
    middle = start + 0.5 * length[:, None] * direction
    oms = range(len(positions)) if candidates is None else candidates
    portions = {}
    for om in oms:
        reach = np.linalg.norm(middle - positions[om], axis=1) <= radius + 0.5 * length
        for index in np.flatnonzero(reach):
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


def cone_harmonics(beta, ratio, wavelengths, phases, degree):
    """Spectral cone harmonics per unit C0, (elements, bands, degree+1)
    (``near_field.py:72-88``): ``sqrt((2l+1)/4pi) P_l(1/(beta n)) (1 + ratio/n^2)/lambda^2``,
    zero where ``beta n <= 1`` or the yield is not positive."""
    beta = np.asarray(beta, float)[:, None]
    ratio = np.asarray(ratio, float)[:, None]
    wavelengths = np.asarray(wavelengths, float)[None, :]
    phases = np.asarray(phases, float)[None, :]
    yield_ = (1 + ratio / phases ** 2) / wavelengths ** 2
    x = 1.0 / (beta * phases)
    legendre = np.empty(x.shape + (degree + 1,))
    legendre[..., 0] = 1.0
    if degree >= 1:
        legendre[..., 1] = x
    for ell in range(1, degree):
        legendre[..., ell + 1] = ((2 * ell + 1) * x * legendre[..., ell]
                                  - ell * legendre[..., ell - 1]) / (ell + 1)
    norm = np.sqrt((2 * np.arange(degree + 1) + 1) / (4 * np.pi))
    out = norm * legendre * yield_[..., None]
    out[((beta * phases <= 1) | (yield_ <= 0))] = 0.0
    return out


def near_field_ge2(kernel, elements, device_pack, time_origin_ns, *, radius_m=None,
                   panel_m=0.05, gauss_order=6, candidates=None, device="cuda"):
    """Clamped order >=2 of every source portion closer than ``radius_m`` (default:
    the pack's radial minimum), relative to ``time_origin_ns``
    (``near_field.py:91-173``). ``device_pack`` is the per-wavelength pack on the
    route's frequencies (a :class:`lighthit_amd.ge2.DevicePack`)."""
    import torch
    low, high = device_pack.radial_low, device_pack.radial_high
    radius = low if radius_m is None else float(radius_m)
    omega = np.asarray(device_pack.omega, float)
    detector = kernel.detector
    positions = np.asarray(detector.positions_m, float)
    n = len(positions)
    origins = np.asarray(time_origin_ns, float)
    if origins.shape != (n,) or not np.isfinite(origins).all():
        raise ValueError("one finite time origin per OM required")
    wave = kernel.wavelength
    wavelengths = np.asarray(wave.wavelength_nm, float)
    phases = np.asarray(kernel.medium.sample(wavelengths)["phase_index"], float)
    efficiency = np.asarray(detector.spectral_weight(wavelengths), float)
    alpha = np.ascontiguousarray(kernel.acceptance()["alpha"], float)
    area = np.broadcast_to(np.asarray(detector.effective_area_m2, float), (n,))
    nodes, weights = np.polynomial.legendre.leggauss(int(gauss_order))
    portions = near_portions(positions, elements, radius, candidates)
    rows = [(om, index, s0, s1) for om in sorted(portions) for index, s0, s1 in portions[om]]
    charge = np.zeros(n)
    spectrum = np.zeros((len(omega), n), complex)
    if rows:
        om_of = np.array([r[0] for r in rows], np.int64)
        index = np.array([r[1] for r in rows], np.int64)
        s0 = np.array([r[2] for r in rows])
        s1 = np.array([r[3] for r in rows])
        axis = np.asarray(elements.direction, float)[index]
        delta = positions[om_of] - np.asarray(elements.start_m, float)[index]
        z = np.einsum("ij,ij->i", delta, axis)
        radial = delta - z[:, None] * axis
        rho = np.linalg.norm(radial, axis=1)
        on_line = rho < 1e-9
        if np.any(on_line):
            # the line goes through the OM centre: any transverse axis
            trial = np.where(np.abs(axis[:, :1]) < 0.9, [[1.0, 0.0, 0.0]], [[0.0, 1.0, 0.0]])
            fallback = trial - np.einsum("ij,ij->i", trial, axis)[:, None] * axis
            radial[on_line] = fallback[on_line]
            rho[on_line] = 1e-9
        ex = radial / np.linalg.norm(radial, axis=1)[:, None]
        ey = np.cross(axis, ex)
        normal = np.asarray(detector.orientations, float)[om_of]
        look = -np.column_stack([np.einsum("ij,ij->i", normal, ex),
                                 np.einsum("ij,ij->i", normal, ey),
                                 np.einsum("ij,ij->i", normal, axis)])
        c0 = np.asarray(elements.coefficient0, float)[index]
        c2 = np.asarray(elements.coefficient2, float)[index]
        ratio = np.where(c0 > 0, c2 / np.where(c0 > 0, c0, 1.0), 0.0)
        harmonics = cone_harmonics(np.asarray(elements.beta, float)[index], ratio, wavelengths,
                                   phases, device_pack.degree)
        keep = np.any(harmonics != 0, axis=(1, 2))
        count = np.maximum(1, np.ceil((s1 - s0) / panel_m).astype(np.int64))
        u_parts, w_parts, node_start = [], [], [0]
        for k in range(len(rows)):
            edges = np.linspace(z[k] - s1[k], z[k] - s0[k], count[k] + 1)
            half = np.diff(edges) / 2
            u_parts.append((edges[:-1, None] + half[:, None] * (nodes + 1)).ravel())
            w_parts.append((half[:, None] * weights).ravel())
            node_start.append(node_start[-1] + (len(u_parts[-1]) if keep[k] else 0))
            if not keep[k]:
                u_parts[-1] = u_parts[-1][:0]
                w_parts[-1] = w_parts[-1][:0]
        length = np.asarray(elements.length_m, float)[index]
        clock = (np.asarray(elements.end_ns, float)[index]
                 - np.asarray(elements.start_ns, float)[index]) / length
        density = c0 / length
        closest = np.asarray(elements.start_ns, float)[index] + z * clock - origins[om_of]
        band_weights = (np.asarray(wave.weight_nm, float)[None, :] * efficiency[None, :]
                        * area[om_of][:, None])

        def up(a, dtype=torch.float64):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)
        keep_alive = dict(impact=up(rho), look=up(look), weight=up(band_weights),
                          harmonics=up(harmonics), phase_ns=up(closest), clock=up(clock),
                          density=up(density), node_start=up(node_start, torch.int32),
                          u=up(np.concatenate(u_parts) if u_parts else np.zeros(1)),
                          w=up(np.concatenate(w_parts) if w_parts else np.zeros(1)))
        desc = _Portions(*[ctypes.c_void_p(keep_alive[k].data_ptr()) for k in
                           ("impact", "look", "weight", "harmonics", "phase_ns", "clock",
                            "density", "node_start", "u", "w")], len(rows))
        out = torch.empty((len(rows), len(omega), 2), dtype=torch.float64, device=device)
        d_alpha = up(alpha[:device_pack.acceptance_degree + 1])
        _lib.call("lh_near_field", ctypes.byref(device_pack.struct), ctypes.byref(desc),
                  _lib.ptr(d_alpha), len(omega), 0.0, float(low), float(high), _lib.ptr(out),
                  _lib.stream_handle())
        parts = out.cpu().numpy()
        parts = parts[..., 0] + 1j * parts[..., 1]
        for k, om in enumerate(om_of):
            if keep[k]:
                spectrum[:, om] += parts[k]
                charge[om] += parts[k, 0].real
    oms = np.array(sorted(portions), int)
    metadata = dict(
        radius_m=radius, table_radial_minimum_m=low, oms=oms.tolist(),
        charge_pe=[float(charge[om]) for om in oms], elements=int(len(rows)),
        approximation="order >=2 Green function clamped at the table radial minimum "
                      "along the same directions; flight time of the actual distance",
        panel_m=float(panel_m), gauss_order=int(gauss_order))
    return NearFieldGe2(oms, charge, spectrum, metadata)


def add_near_field(response, kernel, elements, device_pack, readout, *, label=None,
                   candidates=None, device="cuda", precision="fp32"):
    """Add the clamped near-field order >=2 to a ``HybridTransportResponse``
    (``near_field.py:176-217``): charges of the near OMs, bins of the active ones
    (``readout``, a :class:`lighthit_amd.readout.ReadoutBasis` of the response's
    grids, without demodulation), and the spectrum; recorded in
    ``metadata['near_field']`` and named in a ``RuntimeWarning``."""
    import torch
    near = near_field_ge2(kernel, elements, device_pack, response.time_origin_ns,
                          candidates=candidates, device=device)
    if not len(near.oms):
        return response
    omega = np.asarray(response.omega_per_ns, float)
    if not np.allclose(np.asarray(device_pack.omega, float), omega, rtol=0, atol=1e-12):
        raise ValueError("near-field frequencies differ from the response")
    fine = response.metadata.get("fine_charge_mask")
    fine = np.ones(len(response.active), bool) if fine is None else np.asarray(fine, bool)
    oms = near.oms[fine[near.oms]]
    response.charge_components_pe[oms, 2] += near.charge_pe[oms]
    active = np.asarray(response.active, bool)
    detailed = oms[active[oms]]
    if len(detailed):
        spectrum = torch.as_tensor(np.ascontiguousarray(near.spectrum_pe[:, detailed],
                                                        np.complex64), device=device)
        response.components_pe[detailed, :, 2] += readout.apply(
            spectrum, precision=precision).double().cpu().numpy()
    if response.spectrum_ge2_pe is not None:
        response.spectrum_ge2_pe[:, oms] += (
            near.spectrum_pe[:, oms]
            * np.exp(1j * omega[:, None] * np.asarray(response.time_origin_ns)[None, oms]))
    record = dict(near.metadata, label=label, charge_added_oms=oms.tolist(),
                  bins_added_oms=detailed.tolist())
    response.metadata.setdefault("near_field", []).append(record)
    ids = kernel.detector.identifiers.get("uid") if kernel.detector.identifiers else None
    named = near.oms.tolist() if ids is None else [int(np.asarray(ids)[i]) for i in near.oms]
    warnings.warn(
        f"{len(near.oms)} OMs have source points closer than the {near.metadata['radius_m']:g} m "
        f"radial minimum of the order >=2 tables (OM {'uid' if ids is not None else 'index'} "
        f"{named}); their order >=2 near part uses the near-field clamp "
        f"({sum(near.metadata['charge_pe']):.3g} pe), see metadata['near_field']",
        RuntimeWarning, stacklevel=3)
    return response
