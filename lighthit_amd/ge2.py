"""Stage (b): band-folded order >=2 apply on the GPU.

Host mirror of ``apply_multi_spectral`` (``_shower_basis.py:1126-1198``) and
``apply_folded`` (``shower/ge2.py:99-119``). The folded pack is uploaded once
(:class:`DevicePack`) and stays resident; a source's channels are uploaded
once per event (:class:`DeviceSource`) and every apply of the route (the
predicted and exact zero-frequency charges and the full spectrum) is one
launch of ``lh_ge2_apply``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .axial import channel_layout


class _Pack(ctypes.Structure):
    _fields_ = [("coefficients", ctypes.c_void_p), ("log_r", ctypes.c_void_p),
                ("absorption", ctypes.c_void_p), ("speed", ctypes.c_void_p),
                ("omega", ctypes.c_void_p), ("bands", ctypes.c_int),
                ("intervals", ctypes.c_int), ("degree", ctypes.c_int),
                ("acceptance_degree", ctypes.c_int), ("freq_stride", ctypes.c_int),
                ("flight", ctypes.c_int), ("coefficient_format", ctypes.c_int)]


class _Source(ctypes.Structure):
    _fields_ = [("points_m", ctypes.c_void_p), ("channels_data", ctypes.c_void_p),
                ("reference_ns", ctypes.c_void_p), ("points", ctypes.c_int),
                ("channels", ctypes.c_int), ("degree", ctypes.c_int),
                ("azimuthal_degree", ctypes.c_int), ("freq_stride", ctypes.c_int)]


class _Receivers(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("looks", ctypes.c_void_p),
                ("out_scale", ctypes.c_void_p), ("count", ctypes.c_int)]


_lib.SIGNATURES["lh_ge2_apply"] = (ctypes.c_int, [ctypes.POINTER(_Pack), ctypes.POINTER(_Source),
                                                  ctypes.POINTER(_Receivers), ctypes.c_void_p,
                                                  ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                  ctypes.c_void_p])
_lib.SIGNATURES["lh_ge2_workspace_bytes"] = (ctypes.c_size_t, [ctypes.c_int] * 4)
_lib.SIGNATURES["lh_ge2_apply_workspace_bytes"] = (ctypes.c_size_t, [
    ctypes.POINTER(_Pack), ctypes.POINTER(_Source), ctypes.c_int, ctypes.c_int])

def _workspace(device, nbytes, stream=None):
    """A reusable device workspace of at least ``nbytes`` (one per device and
    stream), or None for none."""
    return _lib.workspace("ge2", device, nbytes, stream) if nbytes else None


_alphas = {}


def _device_alpha(alpha, degree, device):
    """Acceptance coefficients on the device, uploaded once per value."""
    import torch
    alpha = np.ascontiguousarray(np.asarray(alpha, float)[:int(degree) + 1])
    key = (str(device), alpha.tobytes())
    if key not in _alphas:
        _alphas[key] = torch.as_tensor(alpha, device=device)
    return _alphas[key]


def workspace_bytes(bands, freq, receivers, points):
    lib = _lib.load()
    fn = lib.lh_ge2_workspace_bytes
    fn.restype, fn.argtypes = _lib.SIGNATURES["lh_ge2_workspace_bytes"]
    return int(fn(int(bands), int(freq), int(receivers), int(points)))


def apply_workspace_bytes(pack_struct, source_struct, receivers, freq):
    """Device workspace bytes of one ``lh_ge2_apply`` call (either coefficient format)."""
    lib = _lib.load()
    fn = lib.lh_ge2_apply_workspace_bytes
    fn.restype, fn.argtypes = _lib.SIGNATURES["lh_ge2_apply_workspace_bytes"]
    return int(fn(ctypes.byref(pack_struct), ctypes.byref(source_struct), int(receivers),
                  int(freq)))


def split_mfma_layout(coefficients):
    """The pack's coefficients (bands, intervals, degree+1, pairs, 4, frequency) complex64
    as the right operand of the MFMA form: bfloat16 hi + lo (lo the bfloat16 rounding
    of x - hi) laid out (bands, intervals, degree+1, frequency, pair, [re hi, re lo,
    im hi, im lo], power), pairs padded with zeros to a multiple of four."""
    import torch
    x = torch.view_as_real(coefficients)                       # (.., pair, pw, w, 2)
    x = x.permute(0, 1, 2, 5, 3, 6, 4).contiguous()           # (.., w, pair, ri, pw)
    hi = x.to(torch.bfloat16)
    lo = (x - hi.float()).to(torch.bfloat16)
    parts = torch.stack([hi, lo], dim=-2)                      # (.., w, pair, ri, hl, pw)
    npair = parts.shape[4]
    padded = (npair + 3) // 4 * 4
    if padded != npair:
        shape = list(parts.shape)
        shape[4] = padded - npair
        parts = torch.cat([parts, parts.new_zeros(shape)], dim=4)
    return parts.contiguous()


def canonical_pairs(acceptance_degree):
    """(lambda, |mu|) pairs in the order of ``CouplingTable.build`` (``directional.py:164-168``)."""
    lam, mu = [], []
    for a in range(int(acceptance_degree) + 1):
        for m in range(a + 1):
            lam.append(a)
            mu.append(m)
    return np.asarray(lam, np.int64), np.asarray(mu, np.int64)


def _address(tensor):
    return ctypes.c_void_p(tensor.data_ptr()) if tensor is not None else None


class DevicePack:
    """A folded ``Ge2KernelStore``-format bundle resident in device memory."""

    def __init__(self, bundle, device="cuda", coefficient_format="complex64"):
        """``coefficient_format`` "bf16_split" keeps the coefficients as bfloat16 hi + lo
        in the layout of the MFMA form (the precision form's multi-frequency apply)."""
        import torch
        if coefficient_format not in ("complex64", "bf16_split"):
            raise ValueError("coefficient_format must be 'complex64' or 'bf16_split'")
        coefficients = bundle["coefficients"]
        resident = isinstance(coefficients, torch.Tensor)
        if not resident:
            coefficients = np.asarray(coefficients)
        if coefficients.ndim != 6 or coefficients.shape[4] != 4:
            raise ValueError("coefficients must be (bands, intervals, degree+1, pairs, 4, frequency)")
        bands, intervals, ndeg, npair, _, nfreq = coefficients.shape
        self.degree = int(bundle["degree"])
        self.acceptance_degree = int(bundle["acceptance_degree"])
        if ndeg != self.degree + 1:
            raise ValueError("coefficients and degree disagree")
        pl, pm = canonical_pairs(self.acceptance_degree)
        if (not np.array_equal(np.asarray(bundle["pair_lambda"]), pl)
                or not np.array_equal(np.asarray(bundle["pair_mu"]), pm) or npair != len(pl)):
            raise ValueError("pack pairs are not the canonical (lambda, mu) list")
        self.omega = np.asarray(bundle["omega"], float)
        if len(self.omega) != nfreq:
            raise ValueError("omega and coefficients disagree")
        self.log_r = np.asarray(bundle["log_r"], float)
        if len(self.log_r) != intervals + 1:
            raise ValueError("log_r and coefficients disagree")
        self.absorption = np.asarray(bundle["absorption"], float)
        self.speed = np.asarray(bundle["speed"], float)
        self.flight = bundle.get("phase") == "flight"
        self.bands, self.intervals, self.nfreq = bands, intervals, nfreq
        self.radial_low = float(np.exp(self.log_r[0]))
        self.radial_high = float(np.exp(self.log_r[-1]))
        as_dev = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
        if resident:
            self.d_coef = coefficients.to(device=device, dtype=torch.complex64).contiguous()
        else:
            self.d_coef = as_dev(coefficients.astype(np.complex64, copy=False), torch.complex64)
        self.d_log_r = as_dev(self.log_r, torch.float64)
        self.d_absorption = as_dev(self.absorption, torch.float64)
        self.d_speed = as_dev(self.speed, torch.float64)
        self.d_omega = as_dev(self.omega, torch.float64)
        self.coefficient_format = coefficient_format
        if coefficient_format == "bf16_split":
            self.d_coef = split_mfma_layout(self.d_coef)
        self.struct = _Pack(self.d_coef.data_ptr(), self.d_log_r.data_ptr(),
                            self.d_absorption.data_ptr(), self.d_speed.data_ptr(),
                            self.d_omega.data_ptr(), bands, intervals, self.degree,
                            self.acceptance_degree, nfreq, int(self.flight),
                            int(coefficient_format == "bf16_split"))
        self.nbytes = int(self.d_coef.numel() * self.d_coef.element_size())
        self._band_stride = self.nbytes // bands

    def band(self, j):
        """Descriptor of node band ``j`` alone (for sources on different point sets)."""
        s = self.struct
        view = _Pack(s.coefficients + j * self._band_stride, s.log_r, s.absorption + 8 * j,
                     s.speed + 8 * j, s.omega, 1, s.intervals, s.degree, s.acceptance_degree,
                     s.freq_stride, s.flight, s.coefficient_format)
        view.bands = 1
        return view


class DeviceSource:
    """Node-band sources (shared points and frame) resident in device memory."""

    def __init__(self, sources, device="cuda"):
        import torch
        sources = list(sources)
        first = sources[0]
        points = np.asarray(first.points_m(), float)
        for item in sources:
            if (not np.array_equal(item.points_m(), points)
                    or item.channels.shape != first.channels.shape
                    or item.azimuthal_degree != first.azimuthal_degree
                    or not np.array_equal(item.frame.axis, first.frame.axis)):
                raise ValueError("fused spectral sources must share points and angular layout")
        self.frame = first.frame
        self.degree = int(first.degree)
        self.azimuthal_degree = int(first.azimuthal_degree)
        npoints, nchan, nfreq = first.channels.shape
        if nchan != channel_layout(self.degree, self.azimuthal_degree)[-1]:
            raise ValueError("channel count does not match degree and azimuthal degree")
        self.local_points = self.frame.rotate(points - np.asarray(self.frame.centre_m, float))
        channels = np.stack([np.asarray(item.channels) for item in sources]).astype(np.complex64)
        self.reference_ns = np.asarray([float(item.reference_ns) for item in sources])
        self.d_points = torch.as_tensor(np.ascontiguousarray(self.local_points), device=device)
        self.d_channels = torch.as_tensor(np.ascontiguousarray(channels), device=device)
        self.d_reference = torch.as_tensor(self.reference_ns, device=device)
        self.bands, self.points, self.nfreq = len(sources), npoints, nfreq
        self.struct = _Source(self.d_points.data_ptr(), self.d_channels.data_ptr(),
                              self.d_reference.data_ptr(), npoints, nchan, self.degree,
                              self.azimuthal_degree, nfreq)


class DeviceBandSource:
    """Node-band sources compiled on the device (:class:`lighthit_amd.compile.CompiledBand`)
    sharing one point set, in the layout of :class:`DeviceSource`."""

    def __init__(self, bands, frequencies=None):
        import torch
        bands = list(bands)
        first = bands[0]
        self.frame = first.frame
        self.degree = int(first.degree)
        self.azimuthal_degree = int(first.azimuthal_degree)
        npoints, nchan, nfreq = first.channels.shape
        if frequencies is not None:
            nfreq = int(frequencies)
        channels = [b.channels if frequencies is None else b.channels[..., :nfreq] for b in bands]
        self.d_points = first.points.contiguous()
        self.d_channels = torch.stack(channels).contiguous()
        self.reference_ns = np.asarray([float(b.reference_ns) for b in bands])
        self.d_reference = torch.as_tensor(self.reference_ns, device=self.d_points.device)
        self.bands, self.points, self.nfreq = len(bands), npoints, nfreq
        self.local_points = None
        self.struct = _Source(self.d_points.data_ptr(), self.d_channels.data_ptr(),
                              self.d_reference.data_ptr(), npoints, nchan, self.degree,
                              self.azimuthal_degree, nfreq)


class DeviceReceivers:
    """OM positions, looks and detector scales in a source's frame, on the device."""

    def __init__(self, frame, positions, looks, out_scale, device="cuda"):
        import torch
        positions = np.atleast_2d(np.asarray(positions, float))
        centre = np.asarray(frame.centre_m, float)
        self.local_positions = frame.rotate(positions - centre)
        self.local_looks = frame.rotate(np.atleast_2d(np.asarray(looks, float)))
        scale = np.ascontiguousarray(out_scale, float)
        self.count = len(positions)
        self.d_positions, self.d_looks, self.d_scale = _lib.upload(
            [self.local_positions, self.local_looks, scale], device)
        self.struct = _Receivers(self.d_positions.data_ptr(), self.d_looks.data_ptr(),
                                 self.d_scale.data_ptr(), self.count)


def apply_device(pack, source, receivers, alpha, window, *, frequencies=None, out=None, stream=None,
                 band=None):
    """(frequency, OM) complex64 device spectrum of ``source`` at ``receivers``.

    ``frequencies`` counts the leading frequencies evaluated (1: zero frequency);
    ``band`` applies a one-band ``source`` with that node band of ``pack``.
    """
    import torch
    if source.bands != (pack.bands if band is None else 1):
        raise ValueError("one spectral source is needed for each bundled wavelength")
    nfreq = pack.nfreq if frequencies is None else int(frequencies)
    if nfreq != 1 and source.nfreq < nfreq:
        raise ValueError("source and kernel frequency grids differ")
    low, high = map(float, window)
    if (not np.isfinite(low) or not np.isfinite(high) or low < pack.radial_low - 1e-9
            or high > pack.radial_high + 1e-9 or low >= high):
        raise ValueError("radial window must lie inside the kernel cache")
    d_alpha = _device_alpha(alpha, pack.acceptance_degree, source.d_points.device)
    if out is None:
        out = torch.empty((nfreq, receivers.count), dtype=torch.complex64,
                          device=source.d_points.device)
    descriptor = pack.struct if band is None else pack.band(int(band))
    need = apply_workspace_bytes(descriptor, source.struct, receivers.count, nfreq)
    ws = _workspace(source.d_points.device, need, stream) if need else None
    with _lib.stage("ge2_apply" if nfreq > 1 else "ge2_zero"):
        _lib.call("lh_ge2_apply", ctypes.byref(descriptor), ctypes.byref(source.struct),
                  ctypes.byref(receivers.struct), _address(d_alpha), nfreq, low, high,
                  _address(out), _address(ws),
                  ctypes.c_size_t(ws.numel() if ws is not None else 0),
                  _lib.stream_handle(stream))
    return out


def apply_multi_spectral(bundle, sources, receivers_m, look_directions, alpha, out_scale, *,
                         radial_window_m=None, pack=None):
    """Drop-in for ``apply_multi_spectral`` (``_shower_basis.py:1126``); (frequency, OM) complex128."""
    sources = list(sources)
    if not sources or len(sources) != len(bundle["absorption"]):
        raise ValueError("one spectral source is needed for each bundled wavelength")
    pack = DevicePack(bundle) if pack is None else pack
    source = DeviceSource(sources)
    receivers = DeviceReceivers(source.frame, receivers_m, look_directions, out_scale)
    if radial_window_m is None:
        for receiver in receivers.local_positions:
            radius = np.linalg.norm(source.local_points - receiver, axis=1)
            if np.any(radius < pack.radial_low) or np.any(radius > pack.radial_high):
                raise ValueError("Source/receiver distances outside the cache")
        radial_window_m = (pack.radial_low, pack.radial_high)
    out = apply_device(pack, source, receivers, alpha, radial_window_m)
    return out.cpu().numpy().astype(np.complex128)


def apply_folded(folded, sources, positions, looks, alpha, area, window, *, pack=None):
    """Drop-in for ``apply_folded`` (``shower/ge2.py:99-119``)."""
    if not len(positions):
        return np.zeros((len(folded["omega"]), 0), complex)
    out_scale = np.repeat(np.asarray(area, float)[None, :], len(sources), axis=0)
    points = sources[0].points_m()
    if all(np.array_equal(item.points_m(), points) for item in sources[1:]):
        return apply_multi_spectral(folded, sources, positions, looks, alpha, out_scale,
                                    radial_window_m=window, pack=pack)
    pack = DevicePack(folded) if pack is None else pack
    total = 0
    for j, item in enumerate(sources):
        source = DeviceSource([item])
        receivers = DeviceReceivers(source.frame, positions, looks, out_scale[j:j + 1])
        out = apply_device(pack, source, receivers, alpha, window, band=j)
        total = total + out.cpu().numpy().astype(np.complex128)
    return total
