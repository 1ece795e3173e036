"""The shower route: ``ShowerTransport(kernel, data)(shower)`` on the GPU.

Host mirror of ``shower/api.py`` (``ShowerConfig``, ``ShowerTransport``,
``transport_shower``, ``shower/api.py:49-539``), ``shower/ge2.py``
(``node_sources``, ``collapse``, ``delay_groups``), the prompt orders of
``transport_prompt`` that the route uses (``prompt.py:500-947``: first-entry
order 0, grouped order 1 with fused bands and the global screen), the radial
window screens (``shower_basis.py:112-135, 388-439``) and the response type
(``transport.py:197-385``). Every per-element and per-(OM, point) computation
runs in the library: order 0 (``lh_order0_*``), order 1 (``lh_order1_*``),
the source compile (``lh_compile_*``), the order >=2 apply (``lh_ge2_apply``),
the readout (``lh_readout_bins``), the fold (``lh_fold_bundle``) and the
segment bounds (``lh_segment_bounds``). The host keeps the route's control
flow: masks, thresholds, the predictor fallback and the response.
"""
from __future__ import annotations

import ctypes
import json
import threading
import warnings
from collections import OrderedDict
from dataclasses import asdict, dataclass, field, replace
from pathlib import Path
from time import perf_counter, process_time

import numpy as np

from . import _lib
from .axial import AxisFrame, column_bounds
from .model import WavelengthQuadrature, acceptance_alpha, glass_radii, look_directions
from .order1 import PromptSettings

C_M_PER_NS = 0.299792458
#: RTE frequency grid of the prepared tables and the readout grid (``shower/api.py:31-32``).
FREQUENCIES_PER_NS = np.linspace(0.0, 1.2, 65)
READOUT_PER_NS = np.linspace(0.0, 1.2, 321)

_lib.SIGNATURES["lh_segment_bounds"] = (ctypes.c_int, [
    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p])


# --------------------------------------------------------------- configuration

@dataclass(frozen=True)
class RouteConfig:
    """The ``KernelConfig`` fields the shower route reads (``transport.py:40-140``)."""
    omega_per_ns: np.ndarray = field(default_factory=lambda: FREQUENCIES_PER_NS.copy())
    readout_omega_per_ns: np.ndarray | None = field(
        default_factory=lambda: READOUT_PER_NS.copy())
    relative_time_edges_ns: np.ndarray = field(
        default_factory=lambda: np.arange(-60.0, 745.0, 5.0))
    wavelength_nodes: int = 3
    wavelength_range_nm: tuple | None = None
    scattering_degree: int = 8
    source_degree: int = 8
    azimuthal_degree: int = 2
    radial_range_m: tuple = (2.0, 150.0)
    radial_nodes: int = 48
    k_max_per_m: float = 4.0
    k_panel_per_m: float = 0.1
    k_order: int = 6
    threshold_pe: float = 0.01
    detail_threshold_pe: float | None = None

    def __post_init__(self):
        edges = np.asarray(self.relative_time_edges_ns, float)
        if edges.ndim != 1 or len(edges) < 2 or np.any(np.diff(edges) <= 0):
            raise ValueError("relative_time_edges_ns must increase")
        if not self.threshold_pe >= 0:
            raise ValueError("threshold_pe must be nonnegative")
        object.__setattr__(self, "relative_time_edges_ns", edges)
        object.__setattr__(self, "omega_per_ns", np.asarray(self.omega_per_ns, float))


class RouteKernel:
    """Medium, detector and configuration of a route (the parts of
    ``TransportKernel`` the shower route uses)."""

    def __init__(self, medium, detector, config=None):
        self.medium = medium
        self.detector = detector
        self.config = config or RouteConfig()
        low, high = self.config.wavelength_range_nm or medium.wavelength_range_nm
        self.wavelength = WavelengthQuadrature.gauss_legendre(low, high,
                                                              int(self.config.wavelength_nodes))
        self._acceptance = None

    def acceptance(self):
        """Acceptance Legendre coefficients (``transport.py:496-530``)."""
        if self._acceptance is None:
            alpha = acceptance_alpha(self.detector,
                                     scattering_degree=int(self.config.scattering_degree))
            self._acceptance = {"alpha": alpha, "degree": len(alpha) - 1}
        return self._acceptance

    def subset(self, receivers, threshold_pe=None):
        """The kernel on OMs ``receivers`` (``muon/universal_primary.py`` ``subset_kernel``)."""
        return kernel_view(self, receivers=receivers, threshold_pe=threshold_pe)


def kernel_view(kernel, *, receivers=None, threshold_pe=None, config=None):
    """A :class:`RouteKernel` sharing ``kernel``'s medium, wavelength quadrature and
    acceptance, on the OMs ``receivers`` (default all), with ``threshold_pe`` (and no
    detail threshold) or a whole ``config`` replaced. ``kernel`` may be a
    :class:`RouteKernel` or the reference's ``TransportKernel`` (the attributes
    ``medium``, ``detector``, ``config``, ``wavelength`` and ``acceptance()``)."""
    view = RouteKernel.__new__(RouteKernel)
    view.medium = kernel.medium
    view.detector = (kernel.detector if receivers is None
                     else kernel.detector.subset(np.asarray(receivers)))
    view.config = kernel.config if config is None else config
    if threshold_pe is not None:
        view.config = replace(view.config, threshold_pe=float(threshold_pe),
                              detail_threshold_pe=None)
    view.wavelength = kernel.wavelength
    view._acceptance = kernel.acceptance()
    return view


def shower_kernel(kernel, data):
    """The kernel with the route's frequency grids (``shower/api.py:92-112``): 65 table
    frequencies and 321 readout frequencies over 0..1.2 /ns, with the data's
    ``kernel_settings`` applied. The given kernel is returned when it already has
    them; otherwise a view keeps its medium, detector, bins and threshold."""
    config = kernel.config
    settings = dict(getattr(data, "kernel_settings", None) or {})
    if settings:
        config = replace(config, **settings)
    readout = getattr(config, "readout_omega_per_ns", None)
    same = (np.array_equal(np.asarray(config.omega_per_ns), FREQUENCIES_PER_NS)
            and readout is not None and np.array_equal(np.asarray(readout), READOUT_PER_NS))
    if same and not settings:
        return kernel
    forced = dict(omega_per_ns=FREQUENCIES_PER_NS.copy(),
                  readout_omega_per_ns=READOUT_PER_NS.copy())
    rte_cache = getattr(data, "rte_cache", None)
    if hasattr(config, "cache_policy") and rte_cache is not None:
        forced.update(cache_directory=str(rte_cache), cache_policy="require")
    return kernel_view(kernel, config=replace(config, **forced))


@dataclass(frozen=True)
class ShowerBasis:
    """The order >=2 axial basis (``shower_basis.py:30-99``)."""
    cell_m: float = 0.25
    axial_order: int = 3
    transverse: str = "stencil"
    stencil_m: float = 0.2
    stencil_threshold_m: float = 0.03
    time_step_ns: float = 0.02
    radial_high_m: float | None = None
    radial_tail_tolerance_pe: float = 1e-4
    radial_tail_policy: str = "warn"

    def __post_init__(self):
        if self.radial_tail_policy not in ("warn", "raise"):
            raise ValueError("radial_tail_policy must be 'warn' or 'raise'")

    def options(self):
        from .compile import BasisOptions
        return BasisOptions(self.cell_m, self.axial_order, self.transverse, self.stencil_m,
                            self.stencil_threshold_m, self.time_step_ns, "linear")


@dataclass(frozen=True)
class ShowerPrompt(PromptSettings):
    """Orders 0 and 1 of the route (``shower/api.py:36-45``: first-entry direct light,
    fused bands at 96 pixels per face, 192 ring nodes)."""
    max_distance_m: float | None = None
    min_distance_m: float = 0.1


@dataclass(frozen=True)
class ShowerConfig:
    """Numerical choices of the shower route (``shower/api.py:49-94``)."""
    basis: ShowerBasis = field(default_factory=ShowerBasis)
    prompt: ShowerPrompt = field(default_factory=ShowerPrompt)
    frequency_nodes: int = 65
    band_rank: int = 3
    fold_subdivide: int = 1
    predictor_cell_m: float = 2.0
    predictor_safety: float = 2.0
    prompt_delay_ns: float = 20.0
    delayed_cell_ns: float = 20.0
    #: "fp32": the plain form; "mixed": the precision form: the multi-frequency
    #: order >=2 applies as bfloat16 hi + lo products on MFMA and the readout's
    #: inverse Fourier product as FP16 hi + lo on MFMA, FP32 accumulation both;
    #: the zero-frequency charges that select OMs stay FP32
    precision: str = "fp32"

    def __post_init__(self):
        if self.precision not in ("fp32", "mixed"):
            raise ValueError("precision must be 'fp32' or 'mixed'")
        if not 1 <= int(self.band_rank) <= 9 or int(self.fold_subdivide) < 1:
            raise ValueError("band_rank must lie in 1..9 and fold_subdivide be positive")
        if (not self.predictor_safety >= 1 or not self.predictor_cell_m > 0
                or not self.prompt_delay_ns >= 0 or not self.delayed_cell_ns > 0):
            raise ValueError("predictor_safety >= 1 and positive predictor/delay cells required")


# ------------------------------------------------------------------- response

@dataclass
class TransportComponent:
    """One selected signal (``transport.py:197-207``)."""
    charge_pe: np.ndarray
    bins_pe: np.ndarray
    time_origin_ns: np.ndarray
    relative_time_edges_ns: np.ndarray
    active: np.ndarray

    @property
    def rate_pe_per_ns(self):
        return self.bins_pe / np.diff(self.relative_time_edges_ns)[None, :]


@dataclass
class HybridTransportResponse:
    """Physical-time orders 0+1 and frequency-domain order >=2 (``transport.py:332-385``)."""
    detector: object
    omega_per_ns: np.ndarray
    relative_time_edges_ns: np.ndarray
    spectrum_ge2_pe: np.ndarray
    components_pe: np.ndarray
    charge_components_pe: np.ndarray
    time_origin_ns: np.ndarray
    active: np.ndarray
    method: str
    metadata: dict
    computed_orders: tuple = (0, 1, 2)

    @property
    def charge_pe(self):
        return self.charge_components_pe.sum(axis=1)

    @property
    def bins_pe(self):
        return self.components_pe.sum(axis=2)

    def select(self, component="all"):
        choices = {0: (0,), 1: (1,), 2: (2,), "0": (0,), "1": (1,), ">=2": (2,),
                   "0+1": (0, 1), "all": (0, 1, 2)}
        if component not in choices:
            raise ValueError("component must be 0, 1, >=2, 0+1 or all")
        selected = choices[component]
        return TransportComponent(self.charge_components_pe[:, selected].sum(axis=1),
                                  self.components_pe[:, :, selected].sum(axis=2),
                                  self.time_origin_ns, self.relative_time_edges_ns, self.active)

    def save(self, path):
        """``.npz`` with schema ``lighthit/hybrid-response/1``."""
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(
            path, omega_per_ns=self.omega_per_ns,
            relative_time_edges_ns=self.relative_time_edges_ns,
            spectrum_ge2_pe=self.spectrum_ge2_pe, components_pe=self.components_pe,
            charge_components_pe=self.charge_components_pe, time_origin_ns=self.time_origin_ns,
            active=self.active,
            metadata=json.dumps({"schema": "lighthit/hybrid-response/1", "method": self.method,
                                 "computed_orders": list(self.computed_orders), **self.metadata},
                                default=str))
        return path


# -------------------------------------------------------------- small helpers

def segment_bounds(positions, elements, flight=None, device="cuda", chords=None):
    """Distance of every OM to the nearest segment and, with ``flight`` (n_g/c), the
    earliest arrival (``lh_segment_bounds``). ``chords`` (a DeviceChords of these
    elements) saves the upload."""
    import torch

    def up(a):
        return torch.as_tensor(np.ascontiguousarray(a, float), device=device)
    rec = up(positions)
    if chords is not None and chords.host is elements:
        starts, dirs, lengths, t0 = (chords.start_m, chords.direction, chords.length_m,
                                     chords.start_ns)
    else:
        starts, dirs = up(elements.start_m), up(elements.direction)
        lengths, t0 = up(elements.length_m), up(elements.start_ns)
    n = len(positions)
    distance = torch.empty(n, dtype=torch.float64, device=device)
    earliest = torch.empty(n, dtype=torch.float64, device=device) if flight is not None else None
    _lib.call("lh_segment_bounds", _lib.ptr(rec), n, _lib.ptr(starts), _lib.ptr(dirs),
              _lib.ptr(lengths), _lib.ptr(t0), len(elements),
              float(flight) if flight is not None else 0.0, _lib.ptr(distance),
              _lib.ptr(earliest), _lib.stream_handle())
    return distance.cpu().numpy(), (earliest.cpu().numpy() if earliest is not None else None)


def delay_groups(elements, *, prompt_delay_ns, cell_ns, window_end_ns, group_index_max):
    """Prompt / delayed-in-window / beyond-window steps (``shower/ge2.py:122-147``)."""
    c = C_M_PER_NS
    first = int(np.argmin(elements.start_ns))
    distance = np.linalg.norm(elements.start_m - elements.start_m[first], axis=1)
    delay = elements.start_ns - elements.start_ns[first] - distance / c
    prompt = delay <= prompt_delay_ns
    beyond = ~prompt & (delay - (group_index_max - 1.0) * distance / c > window_end_ns)
    within = ~prompt & ~beyond
    groups = []
    if np.any(within):
        cells = np.floor(delay / cell_ns).astype(np.int64)
        for cell in np.unique(cells[within]):
            mask = within & (cells == cell)
            groups.append((mask, float(delay[mask].min())))
    return prompt, groups, beyond


def spectral_attenuation(kernel):
    """Absorption factor of the detected spectrum vs distance as the table
    ``(radius, log_factor)``: factor(d) = exp(interp(d, radius, log_factor))
    (``shower_basis.py:388-405``)."""
    wavelengths = np.asarray(kernel.wavelength.wavelength_nm, float)
    weights = (np.asarray(kernel.wavelength.weight_nm, float)
               * np.asarray(kernel.detector.spectral_weight(wavelengths), float)
               / wavelengths ** 2)
    absorption = np.asarray(kernel.medium.sample(wavelengths)["absorption_per_m"], float)
    lowest = float(absorption.min())
    radius = np.geomspace(1e-3, 1e5, 1025)
    log_factor = (np.log(np.exp(-np.outer(radius, absorption - lowest)) @ weights
                         / weights.sum()) - lowest * radius)
    return radius, log_factor


class _WindowInput(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("channels", ctypes.c_void_p),
                ("count", ctypes.c_int), ("channel_stride", ctypes.c_int),
                ("frame", ctypes.c_void_p), ("positions", ctypes.c_void_p),
                ("areas", ctypes.c_void_p), ("n_oms", ctypes.c_int),
                ("low", ctypes.c_double), ("high", ctypes.c_double),
                ("weight", ctypes.c_double), ("scale", ctypes.c_double),
                ("radius", ctypes.c_void_p), ("log_factor", ctypes.c_void_p),
                ("table", ctypes.c_int)]


_lib.SIGNATURES["lh_window_report"] = (ctypes.c_int, [
    ctypes.POINTER(_WindowInput), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_void_p])


class WindowGeometry:
    """OM positions and areas, the attenuation table and the acceptance peak of
    one kernel on the device, for :func:`window_report`."""

    def __init__(self, kernel, device="cuda"):
        import torch
        detector = kernel.detector
        radius, log_factor = spectral_attenuation(kernel)
        self.peak_acceptance = float(np.max(np.abs(np.asarray(
            detector.angular_acceptance(np.linspace(-1, 1, 257)), float))))
        self.positions = torch.as_tensor(np.ascontiguousarray(detector.positions_m, float),
                                         device=device)
        self.areas = torch.as_tensor(np.ascontiguousarray(detector.effective_area_m2, float),
                                     device=device)
        self.radius = torch.as_tensor(radius, device=device)
        self.log_factor = torch.as_tensor(log_factor, device=device)
        self.count = len(detector.positions_m)
        self.device = device


def window_report(geometry, band, low, high, efficiency, stream=None):
    """Usable, beyond and below-window basis points per OM and the far-tail
    estimate of one compiled band (``shower_basis.py:408-439``), on the device:
    ``(used, far, near, tail)`` tensors over the OMs of ``geometry``."""
    import torch
    n = geometry.count
    device = geometry.device
    used = torch.empty(n, dtype=torch.int32, device=device)
    far = torch.empty(n, dtype=torch.int32, device=device)
    near = torch.empty(n, dtype=torch.int32, device=device)
    tail = torch.empty(n, dtype=torch.float64, device=device)
    points = band.points if band.points.dtype == torch.float64 else band.points.double()
    points = points.contiguous()
    channels = band.channels.contiguous()
    f = band.frame
    frame = torch.as_tensor(np.concatenate([np.asarray(f.centre_m, float), np.asarray(f.axis),
                                            np.asarray(f.first), np.asarray(f.second)]),
                            dtype=torch.float64, device=device)
    stride = int(np.prod(channels.shape[1:])) if channels.dim() > 1 else 1
    w = _WindowInput(_lib.ptr(points), _lib.ptr(channels), int(points.shape[0]), stride,
                     _lib.ptr(frame), _lib.ptr(geometry.positions), _lib.ptr(geometry.areas), n,
                     float(low), float(high), float(np.sqrt(4 * np.pi) / efficiency),
                     float(10 * geometry.peak_acceptance * efficiency),
                     _lib.ptr(geometry.radius), _lib.ptr(geometry.log_factor),
                     int(geometry.radius.numel()))
    with _lib.stage("window_report"):
        _lib.call("lh_window_report", ctypes.byref(w), _lib.ptr(used), _lib.ptr(far),
                  _lib.ptr(near), _lib.ptr(tail), _lib.stream_handle(stream))
    return used, far, near, tail


def screen_radial_tail(tail_pe, basis, high_m):
    """Warn (or raise) for light estimated beyond the cache (``shower_basis.py:112-135``)."""
    tail_pe = np.asarray(tail_pe, float)
    over = np.flatnonzero(tail_pe > basis.radial_tail_tolerance_pe)
    entry = {"policy": basis.radial_tail_policy, "radial_high_m": float(high_m),
             "tolerance_pe": float(basis.radial_tail_tolerance_pe),
             "max_estimate_pe": float(tail_pe.max(initial=0.0)),
             "oms_above_tolerance": over.tolist(), "estimate_pe": tail_pe[over].tolist()}
    if len(over):
        message = (f"{len(over)} OMs have estimated light beyond {high_m:g} m above "
                   f"{basis.radial_tail_tolerance_pe:g} pe; largest estimate "
                   f"{entry['max_estimate_pe']:g} pe")
        if basis.radial_tail_policy == "raise":
            raise ValueError(message)
        warnings.warn(message + " (not computed; listed in metadata['radial_tail'])",
                      RuntimeWarning, stacklevel=3)
    return entry


@dataclass
class CollapsedSource:
    """Zero-frequency source reduced to one point per predictor cell."""
    frame: AxisFrame
    z_m: np.ndarray
    transverse_m: np.ndarray
    channels: np.ndarray
    degree: int
    azimuthal_degree: int
    reference_ns: float

    def points_m(self):
        return (self.frame.centre_m[None, :] + self.z_m[:, None] * self.frame.axis[None, :]
                + self.transverse_m[:, 0:1] * self.frame.first[None, :]
                + self.transverse_m[:, 1:2] * self.frame.second[None, :])


def collapse(bands, cell_m):
    """One point per ``cell_m`` cube of the frame (``shower/ge2.py:70-96``)."""
    reduced = []
    for band in bands:
        points = band.points.cpu().numpy()
        z, transverse = points[:, 2], points[:, :2]
        channels = band.channels[:, :, :1].cpu().numpy().astype(np.complex128)
        weight = np.abs(channels[:, 0, 0])
        keys = np.floor(np.column_stack((z, transverse)) / float(cell_m)).astype(np.int64)
        _, inverse = np.unique(keys, axis=0, return_inverse=True)
        inverse = np.asarray(inverse).ravel()
        chunks = [np.flatnonzero(inverse == g) for g in range(int(inverse.max()) + 1)]
        zz, tt, cc = [], [], []
        for chunk in chunks:
            total = weight[chunk].sum()
            share = (weight[chunk] / total if total > 0 else np.full(len(chunk), 1.0 / len(chunk)))
            zz.append(np.sum(share * z[chunk]))
            tt.append(np.sum(share[:, None] * transverse[chunk], axis=0))
            cc.append(channels[chunk].sum(axis=0))
        reduced.append(CollapsedSource(band.frame, np.asarray(zz), np.asarray(tt),
                                       np.ascontiguousarray(np.asarray(cc)), band.degree,
                                       band.azimuthal_degree, band.reference_ns))
    return reduced


# ---------------------------------------------------------------- order >= 2

def _same_points(bands):
    import torch
    first = bands[0].points
    return all(b.points.shape == first.shape and torch.equal(b.points, first) for b in bands[1:])


def apply_bands(pack, bands, positions, looks, alpha, area, window, *, frequencies=None):
    """Folded order >=2 spectrum (frequency, OM) complex128 of node-band sources
    (device :class:`CompiledBand` or host sources), ``apply_folded`` semantics."""
    from .ge2 import DeviceBandSource, DeviceReceivers, DeviceSource, apply_device
    nfreq = pack.nfreq if frequencies is None else int(frequencies)
    if not len(positions):
        return np.zeros((nfreq, 0), complex)
    area = np.asarray(area, float)
    device_bands = hasattr(bands[0], "points") and not callable(getattr(bands[0], "points"))
    if device_bands and _same_points(bands):
        source = DeviceBandSource(bands, frequencies)
        receivers = DeviceReceivers(source.frame, positions, looks,
                                    np.repeat(area[None, :], len(bands), 0))
        out = apply_device(pack, source, receivers, alpha, window, frequencies=frequencies)
        return out.cpu().numpy().astype(np.complex128)
    if not device_bands and all(np.array_equal(b.points_m(), bands[0].points_m())
                                for b in bands[1:]):
        source = DeviceSource(bands)
        receivers = DeviceReceivers(source.frame, positions, looks,
                                    np.repeat(area[None, :], len(bands), 0))
        out = apply_device(pack, source, receivers, alpha, window, frequencies=frequencies)
        return out.cpu().numpy().astype(np.complex128)
    total = 0
    for j, band in enumerate(bands):
        source = DeviceBandSource([band], frequencies) if device_bands else DeviceSource([band])
        receivers = DeviceReceivers(source.frame, positions, looks, area[None, :])
        out = apply_device(pack, source, receivers, alpha, window, frequencies=frequencies, band=j)
        total = total + out.cpu().numpy().astype(np.complex128)
    return total


def node_sources(kernel, chords, basis, fold, omega, frame, mask=None, omega_device=None):
    """Node-band sources compiled on the device, scaled by ``fold.scale``
    (``shower/ge2.py:35-67``)."""
    import torch
    from .compile import CompiledBand, compile_band
    wavelength = np.asarray(kernel.wavelength.wavelength_nm, float)
    compiled = {}
    for band in fold.bands:
        item = compile_band(chords, frame, int(kernel.config.source_degree), omega,
                            azimuthal_degree=int(kernel.config.azimuthal_degree),
                            options=basis.options(), wavelength_nm=float(wavelength[band]),
                            phase_index=float(fold.phase_index[band]), mask=mask,
                            scale=float(fold.scale[band]), omega_device=omega_device)
        if item is not None:
            compiled[band] = item
    if not compiled:
        raise ValueError("no element emits in any node band")
    template = next(iter(compiled.values()))
    return [compiled.get(band) or CompiledBand(template.frame, template.points,
                                               torch.zeros_like(template.channels),
                                               template.degree, template.azimuthal_degree,
                                               template.reference_ns, dict(template.summary))
            for band in fold.bands]


# ------------------------------------------------------------------- prompt

@dataclass
class PromptResult:
    """Orders 0 and 1 of the route (the fields of ``PromptTransportResponse`` it reads)."""
    charge_orders_pe: np.ndarray      # (N, 2)
    bins_orders_pe: np.ndarray        # (N, nb, 2)
    active: np.ndarray
    computed: np.ndarray
    time_origin_ns: np.ndarray
    level: np.ndarray
    metadata: dict

    @property
    def charge_pe(self):
        return self.charge_orders_pe.sum(axis=1)


def transport_prompt(kernel, elements, settings, *, order0=None, selection_extra_charge=None,
                     time_origin_ns=None, compute_bins_for_all=False, device="cuda",
                     chords=None):
    """Orders 0 and 1 of a spectral-cone element source (``prompt.py:500-947`` with
    the shower route's settings: first-entry order 0, grouped fused order 1)."""
    from .ballistic import DeviceElements, Order0
    from .compile import DeviceChords
    from .order1 import DeviceOMs, spectral_order1, spectral_tables
    began = perf_counter()
    timings = {}
    if compute_bins_for_all and settings.screen:
        settings = replace(settings, screen=False)
    tables = spectral_tables(kernel.medium, kernel.detector, kernel.wavelength)
    alpha = np.ascontiguousarray(kernel.acceptance()["alpha"], float)
    detector = kernel.detector
    positions = np.ascontiguousarray(detector.positions_m, float)
    n = len(positions)
    extra = np.zeros(n) if selection_extra_charge is None else np.asarray(selection_extra_charge,
                                                                          float)
    if extra.shape != (n,) or not np.isfinite(extra).all():
        raise ValueError("selection_extra_charge must be finite per OM")
    looks = look_directions(detector)
    areas = np.asarray(detector.effective_area_m2, float)
    edges = np.asarray(kernel.config.relative_time_edges_ns, float)
    threshold = float(kernel.config.threshold_pe)
    g = float(kernel.medium.g)
    # threshold of the continuous spectrum (prompt.py:543-552)
    low, high = kernel.config.wavelength_range_nm or kernel.medium.wavelength_range_nm
    native = kernel.medium.wavelength_nm
    critical = np.r_[low, native[(native > low) & (native < high)], high]
    threshold_phase = float(np.max(kernel.medium.sample(critical)["phase_index"]))
    valid = elements.beta * threshold_phase > 1
    if not np.any(valid):
        raise ValueError("source is below Cherenkov threshold over the wavelength range")
    if not np.all(valid):
        elements = elements.subset(valid)
    fastest = float(np.min(tables["group_index"]))
    if time_origin_ns is not None:
        origins = np.asarray(time_origin_ns, float)
        if origins.shape != (n,) or not np.isfinite(origins).all():
            raise ValueError("time_origin_ns must be finite with one value per OM")
    else:
        first = int(np.argmin(elements.start_ns))
        origins = (float(elements.start_ns[first])
                   + np.linalg.norm(positions - elements.start_m[first], axis=1)
                   / (C_M_PER_NS / fastest))
    # module selection (prompt.py:596-626)
    stage = perf_counter()
    max_distance = (float(kernel.config.radial_range_m[1]) if settings.max_distance_m is None
                    else float(settings.max_distance_m))
    if chords is None or chords.host is not elements:
        chords = DeviceChords(elements, device)
    distance, _ = segment_bounds(positions, elements, None, device, chords)
    inside = distance <= max_distance
    density = (elements.coefficient0[:, None] / tables["wavelength_nm"][None, :] ** 2
               + elements.coefficient2[:, None]
               / (tables["wavelength_nm"][None, :] ** 2 * tables["phase_index"][None, :] ** 2))
    emitted = float(np.sum(tables["weight_nm"][None, :] * np.maximum(density, 0.0)))
    peak_acceptance = float(np.max(np.abs(np.asarray(
        detector.angular_acceptance(np.linspace(-1, 1, 257)), float))))
    far = np.maximum(distance, 1e-9)
    upper = (10 * max(emitted, 0.0) * areas * peak_acceptance
             * float(np.max(tables["efficiency"]))
             * np.exp(-float(np.min(tables["mua"])) * far) / (4 * np.pi * far ** 2))
    unsafe = (~inside) & (upper >= threshold)
    if np.any(unsafe):
        raise ValueError(f"{int(unsafe.sum())} OMs beyond max_distance_m={max_distance:g} have a "
                         "conservative estimate above threshold; increase max_distance_m")
    too_close = inside & (distance < settings.min_distance_m)
    if np.any(too_close):
        raise ValueError(f"{int(too_close.sum())} OMs are closer than min_distance_m="
                         f"{settings.min_distance_m:g} m to the source")
    work = np.flatnonzero(inside)
    timings["selection_seconds"] = perf_counter() - stage
    # order 0 (first entry)
    stage = perf_counter()
    if order0 is None:
        order0 = Order0(kernel.medium, detector, alpha, device=device)
    device_elements = DeviceElements.from_chords(chords)
    charge = np.zeros((n, 2))
    if len(work):
        # the OMs of the work set only, as the reference evaluates its subset
        # detector (prompt.py:641-650); every OM's charge is the same either way
        charge[work, 0] = order0.charge(device_elements, ids=work)
    timings["order0_charge_seconds"] = perf_counter() - stage
    # order 1
    stage = perf_counter()
    oms = DeviceOMs(positions, looks, areas, origins, device)
    q1, bins1, level = spectral_order1(chords, oms, work, tables, alpha, g, edges, threshold,
                                       charge[:, 0], selection_extra_charge, settings)
    charge[:, 1] = q1
    timings["order1_seconds"] = perf_counter() - stage
    computed = inside.copy()
    active = computed & (charge.sum(axis=1) + extra >= threshold)
    if threshold == 0:
        active = computed.copy()
    if selection_extra_charge is not None and np.any(active & (level < 2)):
        raise ValueError("prompt screen omitted an OM selected by the combined charge")
    stage = perf_counter()
    bins = np.zeros((n, len(edges) - 1, 2))
    index = work if compute_bins_for_all else np.flatnonzero(active)
    if len(index):
        bins[index, :, 0] = order0.bins(device_elements, index, origins[index], edges)
        bins[index, :, 1] = bins1[index]
    timings["order0_bins_seconds"] = perf_counter() - stage
    timings["elapsed_seconds"] = perf_counter() - began
    metadata = {"computed_orders": [0, 1], "prompt_algorithm": "grouped",
                "max_distance_m": max_distance, "computed_modules": int(computed.sum()),
                "active_modules": int(active.sum()), "timings": timings,
                "order1_level": level.tolist(),
                "ballistic": "first-entry", "source_elements": len(elements)}
    return PromptResult(charge, bins, active, computed, origins, level, metadata)


# -------------------------------------------------------------------- route

MANIFEST = "shower-data.json"
ENVIRONMENT = "LIGHTHIT_SHOWER_DATA"


@dataclass(frozen=True)
class ShowerData:
    """Tables of the route (``shower/data.py:26-100``): the RTE cache directory, the
    order >=2 pack directory (or an in-memory :class:`KernelPack`), an optional
    prepared fold, the manifest and its ``kernel_settings``. The route reads the
    pack (and the fold); ``rte_cache`` is carried for the kernel's configuration
    only, since the RTE tables enter the route through the pack."""
    rte_cache: object
    ge2_pack: object
    ge2_band_fold: object = None
    manifest: Path | None = None
    kernel_settings: dict = field(default_factory=dict)

    @classmethod
    def of(cls, rte_cache, ge2_pack, ge2_band_fold=None):
        """Tables given directly; paths must exist (``rte_cache`` may be None)."""
        def resolve(item):
            return item if item is None or not isinstance(item, (str, Path)) else \
                Path(item).expanduser().resolve()
        cache, pack = resolve(rte_cache), resolve(ge2_pack)
        absent = [str(item) for item in (cache, pack) if isinstance(item, Path)
                  and not item.exists()]
        if absent:
            raise FileNotFoundError("shower data incomplete: " + ", ".join(absent))
        return cls(cache, pack, resolve(ge2_band_fold))

    @classmethod
    def load(cls, path):
        """Read ``shower-data.json`` (or ``muon-data.json``) or a directory holding one."""
        path = Path(path).expanduser()
        if path.is_dir():
            path = path / MANIFEST if (path / MANIFEST).is_file() else path / "muon-data.json"
        entries = json.loads(path.read_text())
        if "ge2_pack" not in entries:
            raise ValueError(f"{path}: missing entries ['ge2_pack']")
        base = path.parent
        cache, fold = entries.get("rte_cache"), entries.get("ge2_band_fold")
        data = cls(None if cache is None else (base / cache).resolve(),
                   (base / entries["ge2_pack"]).resolve(),
                   None if fold is None else (base / fold).resolve(), path.resolve(),
                   dict(entries.get("kernel_settings") or {}))
        absent = [str(item) for item in (data.rte_cache, data.ge2_pack)
                  if item is not None and not item.exists()]
        if absent:
            raise FileNotFoundError("shower data incomplete: " + ", ".join(absent))
        return data

    @classmethod
    def find(cls, cache_directory=None):
        """``$LIGHTHIT_SHOWER_DATA``, then a shower (or muon) manifest in or next to the
        RTE cache directory (``shower/data.py:84-100``)."""
        import os
        places = []
        if os.environ.get(ENVIRONMENT):
            places.append(Path(os.environ[ENVIRONMENT]))
        if cache_directory is not None:
            cache = Path(cache_directory).expanduser()
            for name in (MANIFEST, "muon-data.json"):
                places += [cache / name, cache.parent / name]
        for place in places:
            if place.is_file() or (place.is_dir() and ((place / MANIFEST).is_file()
                                                       or (place / "muon-data.json").is_file())):
                return cls.load(place)
        searched = ", ".join(str(item) for item in places) or "nothing (no cache directory)"
        raise FileNotFoundError(f"no {MANIFEST} or muon-data.json found; searched {searched}")


class ShowerTransport:
    """Transport for shower events on one detector (``shower/api.py:115-220``).

    ``kernel`` is a :class:`RouteKernel` or the reference's ``TransportKernel``;
    ``data`` is a :class:`ShowerData`, a manifest path, a
    :class:`lighthit_amd.fold.KernelPack`, or None to search next to the
    kernel's cache (:meth:`ShowerData.find`). The kernel is given the route's
    frequency grids (:func:`shower_kernel`). The pack is folded onto the node
    bands once (on the GPU) and stays resident with the order-0 optics and
    the readout basis.
    """

    def __init__(self, kernel, data=None, *, config=None, threads=None, device="cuda"):
        from .fold import KernelPack
        if data is None:
            data = ShowerData.find(getattr(kernel.config, "cache_directory", None))
        elif isinstance(data, KernelPack):
            data = ShowerData(None, data)
        elif not isinstance(data, ShowerData):
            data = ShowerData.load(data)
        self.data = data
        self.kernel = shower_kernel(kernel, data)
        self.config = config or ShowerConfig()
        self.threads = threads
        self.device = device
        self._tables = None
        self._local = threading.local()   # per-thread cache of the last compile
        self._lock = threading.RLock()     # guards the lazily built resident parts
        self._order0 = None
        self._readout = OrderedDict()
        self._window = None

    @property
    def _compiled(self):
        cache = getattr(self._local, "compiled", None)
        if cache is None:
            cache = self._local.compiled = {}
        return cache

    def tables(self):
        """The pack folded onto the node bands, resident on the device (once)."""
        with self._lock:
            return self._tables_locked()

    def _tables_locked(self):
        if self._tables is None:
            from .fold import BandFold, KernelPack, fold_bundle, frequency_indices, load_folded
            from .ge2 import DevicePack
            pack = self.data.ge2_pack
            if not isinstance(pack, KernelPack):
                pack = KernelPack.load(pack)
            kernel = self.kernel
            rank, subdivide = int(self.config.band_rank), int(self.config.fold_subdivide)
            indices = frequency_indices(kernel.config.omega_per_ns, self.config.frequency_nodes)
            omega = kernel.config.omega_per_ns[indices]
            folded = None
            prepared = self.data.ge2_band_fold
            if prepared is not None:
                prepared = load_folded(prepared) if not isinstance(prepared, dict) else prepared
                if (prepared["band_fold"]["rank"] == rank
                        and int(prepared.get("subdivide", 1)) == subdivide
                        and np.asarray(prepared["omega"]).shape == omega.shape
                        and np.allclose(prepared["omega"], omega, rtol=0, atol=1e-12)):
                    folded = prepared
            if folded is None:
                fold = BandFold.of(kernel, rank=rank)
            else:
                fold = BandFold.of(kernel, rank=rank, bands=folded["band_fold"]["bands"])
                stored = folded["band_fold"]["worst_residual_design"]
                if not np.isclose(fold.metadata()["worst_residual_design"], stored, rtol=1e-9,
                                  atol=0):
                    raise ValueError("the prepared fold was made for another medium or detector")
            began = perf_counter()
            if folded is None:
                bundle = pack.bundle(kernel.wavelength.wavelength_nm,
                                     int(kernel.config.source_degree), indices)
                folded = fold_bundle(bundle, fold, subdivide=subdivide, device=self.device)
            device_pack = DevicePack(folded, self.device)
            self._tables = dict(pack=pack, fold=fold, folded=folded, device_pack=device_pack,
                                fold_seconds=perf_counter() - began)
            if self.config.precision == "mixed":
                # the multi-frequency applies of the precision form (MFMA layout)
                self._tables["device_pack_split"] = DevicePack(folded, self.device,
                                                               coefficient_format="bf16_split")
        return self._tables

    def order0(self):
        with self._lock:
            if self._order0 is None:
                from .ballistic import Order0
                self._order0 = Order0(self.kernel.medium, self.kernel.detector,
                                      self.kernel.acceptance()["alpha"], device=self.device)
            return self._order0

    def near_pack(self, kernel):
        """The per-wavelength pack of the near-field clamp (built on first use)."""
        with self._lock:
            tables = self.tables()
            if tables.get("near_pack") is None:
                tables["near_pack"] = _near_pack(tables, kernel, self.config, self.device)
            return tables["near_pack"]

    def window_geometry(self, kernel):
        """Device geometry of the radial-window report (once per detector)."""
        with self._lock:
            if self._window is None or self._window[0] is not kernel.detector:
                self._window = (kernel.detector, WindowGeometry(kernel, self.device))
            return self._window[1]

    #: readout bases a route keeps (the prompt grid and recent delayed-light shifts)
    READOUT_CACHE = 8

    def readout(self, omega, edges):
        """The readout basis of (omega, edges, readout grid), kept per route in a
        small least-recently-used cache (delayed light reads on shifted edges)."""
        return _cached_readout(self._readout, self._lock, self.kernel, omega, edges,
                               self.device, self.READOUT_CACHE)

    def __call__(self, source, *, progress=None, time_origin_ns=None, selection_extra_charge=None,
                 receivers=None):
        """Response for one event (``HybridTransportResponse``)."""
        return self._call(source, time_origin_ns=time_origin_ns,
                          selection_extra_charge=selection_extra_charge, receivers=receivers)

    def batch(self, sources, *, graph=False, **call_kwargs):
        """Responses of several events (the serving form's batch): the order-0
        charges of all of them in one fused persistent launch
        (``lh_order0_charge_batch``; with ``graph`` replayed from a HIP graph
        captured once per capacity), the rest of each event as in a plain call.
        Each response equals the plain call's bit for bit; an event that
        raises gives its exception in the returned list in place of a response.
        ``call_kwargs`` (``time_origin_ns``, ``selection_extra_charge``) apply
        to every event; a ``receivers`` subset is not batched."""
        import torch
        from .ballistic import DeviceElements, check_flags
        if call_kwargs.get("receivers") is not None:
            raise ValueError("batched events take the route's whole detector")
        residents = []
        for source in sources:
            try:
                residents.append(source if isinstance(source, ResidentShower)
                                 else ResidentShower(source, self.device))
            except (TypeError, ValueError) as error:
                residents.append(error)
        fused = []
        for resident in residents:
            if isinstance(resident, ResidentShower):
                try:
                    fused.append(DeviceElements.from_chords(resident.chords))
                except ValueError:
                    pass        # the plain call raises it for this event
        order0 = self.order0()
        prefetched = {}
        # the OMs any event's order 0 evaluates: within the prompt's selection
        # radius of its elements (the work sets of transport_prompt)
        positions = np.ascontiguousarray(self.kernel.detector.positions_m, float)
        reach = (float(self.kernel.config.radial_range_m[1])
                 if self.config.prompt.max_distance_m is None
                 else float(self.config.prompt.max_distance_m))
        union = np.zeros(len(positions), bool)
        for elements in fused:
            chords = elements.chords
            distance, _ = segment_bounds(positions, chords.host, None, self.device, chords)
            union |= distance <= reach
        ids = np.flatnonzero(union)
        if fused and len(ids):
            out, ws = order0.charge_batch_device(fused, graph=graph and len(ids) == len(union),
                                                 oms=None if len(ids) == len(union)
                                                 else order0.oms_of(ids))
            torch.cuda.current_stream().synchronize()
            try:
                check_flags(ws)
                host = out.cpu().numpy()
                for row, elements in enumerate(fused):
                    prefetched[(elements.struct.start_m, elements.count)] = host[row]
            except ValueError:
                pass            # the event at fault raises in its own call
        order0 = _PrefetchedOrder0(order0, prefetched, ids)
        responses = []
        for resident in residents:
            if not isinstance(resident, ResidentShower):
                responses.append(resident)
                continue
            try:
                responses.append(self._call(resident, order0=order0, **call_kwargs))
            except (ValueError, RuntimeError, TypeError) as error:
                responses.append(error)
        return responses

    def _call(self, source, *, time_origin_ns=None, selection_extra_charge=None, receivers=None,
              order0=None):
        # detail threshold (transport.py:210-254): every selection decision at the
        # lowered threshold, ``active`` at the original one -- also for a receiver
        # subset, whose own threshold is zero
        kernel = self.kernel
        detail = getattr(kernel.config, "detail_threshold_pe", None)
        lowered = detail is not None and detail < kernel.config.threshold_pe
        if lowered:
            kernel = kernel_view(kernel, threshold_pe=float(detail))
        order0 = self.order0() if order0 is None else order0
        if receivers is not None:
            kernel = kernel_view(kernel, receivers=receivers, threshold_pe=0.0)
            order0 = None
        response = transport_shower(kernel, source, config=self.config, tables=self.tables(),
                                    time_origin_ns=time_origin_ns,
                                    selection_extra_charge=selection_extra_charge,
                                    compiled=self._compiled, route=self if receivers is None
                                    else None, order0=order0, device=self.device)
        response.metadata["route"] = "lighthit.shower"
        response.metadata["fold_prepare_seconds"] = self.tables()["fold_seconds"]
        if lowered:
            threshold = float(self.kernel.config.threshold_pe)
            charge = response.charge_pe + (0 if selection_extra_charge is None
                                           else np.asarray(selection_extra_charge, float))
            response.active = response.active & (charge >= threshold)
            response.metadata["threshold_pe"] = threshold
            response.metadata["detail_threshold_pe"] = float(detail)
            response.metadata["active_modules"] = int(np.count_nonzero(response.active))
        return response


class _PrefetchedOrder0:
    """The route's order-0 evaluator with the charges of some events already
    computed (by one fused launch on a set of OMs): ``charge`` of those
    events' device elements on OMs of that set returns them, everything else
    goes to the evaluator itself."""

    def __init__(self, order0, prefetched, ids):
        self._order0 = order0
        self._prefetched = prefetched
        self._ids = ids

    def charge(self, elements, ids=None):
        struct = getattr(elements, "struct", None)
        got = None if struct is None else self._prefetched.get((struct.start_m, elements.count))
        if got is not None:
            want = np.arange(len(self._order0.detector.positions_m)) if ids is None \
                else np.asarray(ids)
            at = np.searchsorted(self._ids, want)
            if np.all(at < len(self._ids)) and np.array_equal(self._ids[np.minimum(
                    at, len(self._ids) - 1)], want):
                return got[at].copy()
        return self._order0.charge(elements, ids=ids)

    def __getattr__(self, name):
        return getattr(self._order0, name)


class ResidentShower:
    """A shower source whose element arrays are already on the device.

    ``ResidentShower(source)`` uploads the elements once; handing it to
    :class:`ShowerTransport` skips that copy in the call (the route otherwise
    uploads the host arrays of every source it is given). It carries the
    host elements too, for the host-side steps of the route.
    """

    def __init__(self, source, device="cuda"):
        from .compile import DeviceChords
        self.elements = _elements_of(source)
        self.chords = DeviceChords(self.elements, device)


def _elements_of(source):
    elements = getattr(source, "elements", source)
    if not hasattr(elements, "coefficient0"):
        raise TypeError("the shower route takes spectral light elements or a synthetic shower")
    return elements


def transport_shower(kernel, source, *, config=None, tables, time_origin_ns=None,
                     selection_extra_charge=None, compiled=None, route=None, order0=None,
                     device="cuda"):
    """All orders for one shower source (``shower/api.py:223-539``)."""
    import torch
    from .compile import DeviceChords
    from .fold import frequency_indices
    config = config or ShowerConfig()
    basis = config.basis
    readout_precision = "fp32" if config.precision == "fp32" else "split"
    began, cpu = perf_counter(), process_time()
    timings = {}
    stage = perf_counter()
    indices = frequency_indices(kernel.config.omega_per_ns, config.frequency_nodes)
    omega = kernel.config.omega_per_ns[indices]
    if omega[0] != 0.0:
        raise ValueError("the shower route needs omega = 0 in its frequency set")
    output_omega = (kernel.config.readout_omega_per_ns
                    if kernel.config.readout_omega_per_ns is not None else omega)
    elements = _elements_of(source)
    if elements.cone_model != "spectral":
        raise ValueError("the shower route requires spectral cones (cone_model='spectral')")
    wavelengths = np.asarray(kernel.wavelength.wavelength_nm, float)
    sample = kernel.medium.sample(wavelengths)
    max_phase = float(np.max(sample["phase_index"]))
    valid = elements.beta * max_phase > 1
    if not np.any(valid):
        raise ValueError("source is below Cherenkov threshold over the wavelength range")
    original_coefficient = float(elements.coefficient0.sum())
    if not np.all(valid):
        elements = elements.subset(valid)
    efficiency = np.asarray(kernel.detector.spectral_weight(wavelengths), float)
    positions = np.asarray(kernel.detector.positions_m, float)
    looks = look_directions(kernel.detector)
    area = np.asarray(kernel.detector.effective_area_m2, float)
    alpha = np.ascontiguousarray(kernel.acceptance()["alpha"], float)
    low, cache_high = kernel.config.radial_range_m
    high = cache_high if basis.radial_high_m is None else basis.radial_high_m
    threshold = float(kernel.config.threshold_pe)
    n = len(positions)
    extra = (np.zeros(n) if selection_extra_charge is None
             else np.maximum(np.asarray(selection_extra_charge, float), 0.0))
    if extra.shape != (n,) or not np.isfinite(extra).all():
        raise ValueError("selection_extra_charge must hold one finite value per OM")
    fold, folded, pack = tables["fold"], tables["folded"], tables["device_pack"]
    full_pack = pack
    if config.precision == "mixed":
        from .ge2 import DevicePack
        full_pack = tables.get("device_pack_split") or DevicePack(
            folded, device, coefficient_format="bf16_split")
    if not np.allclose(folded["omega"], omega, rtol=0, atol=1e-12):
        raise ValueError("folded >=2 kernels and route frequencies differ")
    timings["setup_seconds"] = perf_counter() - stage

    # order >=2 source on the node bands
    stage = perf_counter()
    edges = np.asarray(kernel.config.relative_time_edges_ns, float)
    prompt_mask, late_groups, beyond = delay_groups(
        elements, prompt_delay_ns=float(config.prompt_delay_ns),
        cell_ns=float(config.delayed_cell_ns), window_end_ns=float(edges[-1]),
        group_index_max=float(np.max(sample["group_index"])))
    key = (fold.bands, basis, tuple(np.round(omega, 12)), int(kernel.config.source_degree),
           int(kernel.config.azimuthal_degree), float(config.prompt_delay_ns),
           float(config.delayed_cell_ns), float(edges[-1]))
    resident = getattr(source, "chords", None)
    if compiled is not None and compiled.get("source") is source and compiled.get("key") == key:
        sources, late_zero, late_sources = compiled["sources"]
        chords = compiled["chords"]
    else:
        chords = (resident if resident is not None and resident.host is elements
                  else DeviceChords(elements, device))
        omega_device = torch.as_tensor(omega, device=device)

        def compiled_part(mask, frequencies):
            part = elements if np.all(mask) else elements.subset(mask)
            frame = AxisFrame.of(part.field(0, phase_index=max_phase))
            return node_sources(kernel, chords, basis, fold, frequencies, frame,
                                mask=None if np.all(mask) else mask,
                                omega_device=omega_device if len(frequencies) == len(omega)
                                else None)
        sources = compiled_part(prompt_mask, omega)
        late_zero = compiled_part(~prompt_mask, omega[:1]) if np.any(~prompt_mask) else None
        late_sources = [(shift, compiled_part(mask, omega)) for mask, shift in late_groups]
        if compiled is not None:
            compiled.clear()
            compiled.update(source=source, key=key, sources=(sources, late_zero, late_sources),
                            chords=chords)
    timings["compile_seconds"] = perf_counter() - stage

    # radial windows of every node-band point set
    stage = perf_counter()
    geometry = (route.window_geometry(kernel) if route is not None
                else WindowGeometry(kernel, device))
    used_dev = torch.zeros(n, dtype=torch.int32, device=device)
    far_dev = torch.zeros(n, dtype=torch.int32, device=device)
    near_dev = torch.zeros(n, dtype=torch.int32, device=device)
    tail_dev = torch.zeros(n, dtype=torch.float64, device=device)
    weight_ratio = float(np.sum(fold.scale) / np.sum(fold.scale[list(fold.bands)]))
    for group in [sources] + ([late_zero] if late_zero is not None else []):
        for band, item in zip(fold.bands, group):
            part_used, part_far, part_near, part_tail = window_report(
                geometry, item, low, high, float(efficiency[band]))
            near_dev = torch.maximum(near_dev, part_near)
            used_dev = torch.maximum(used_dev, part_used)
            far_dev = torch.maximum(far_dev, part_far)
            tail_dev += weight_ratio * part_tail
    used = used_dev.cpu().numpy().astype(np.int64)
    far = far_dev.cpu().numpy().astype(np.int64)
    near = near_dev.cpu().numpy() > 0
    tail = tail_dev.cpu().numpy()
    radial_tail = screen_radial_tail(tail, basis, high)
    reach = np.flatnonzero(used > 0)

    # predicted zero-frequency >=2 charge of every reachable OM
    predicted = np.zeros(n)
    collapsed = collapse(sources, float(config.predictor_cell_m))
    late_charge = np.zeros(n)
    if len(reach):
        predicted[reach] = apply_bands(pack, collapsed, positions[reach], looks[reach], alpha,
                                       area[reach], (low, high), frequencies=1)[0].real
        if late_zero is not None:
            late_charge[reach] = apply_bands(pack, late_zero, positions[reach], looks[reach],
                                             alpha, area[reach], (low, high),
                                             frequencies=1)[0].real
    timings["predict_seconds"] = perf_counter() - stage

    # time origins (shower/api.py:356-375)
    fastest = float(np.min(sample["group_index"]))
    time_width = float(np.ptp(edges))
    source_time_span = float(np.max(elements.end_ns) - np.min(elements.start_ns))
    low_m, high_m = column_bounds(elements.midpoints_m)
    source_extent = float(np.linalg.norm(high_m - low_m))
    local_origin = source_extent > high or source_time_span > 0.5 * time_width
    if time_origin_ns is not None:
        origins = np.asarray(time_origin_ns, float)
        if origins.shape != (n,) or not np.isfinite(origins).all():
            raise ValueError("time_origin_ns must hold one finite origin per OM")
        local_origin = True
    elif local_origin:
        _, origins = segment_bounds(positions, elements, fastest / C_M_PER_NS, device, chords)
    else:
        first = int(np.argmin(elements.start_ns))
        origins = (float(elements.start_ns[first])
                   + np.linalg.norm(positions - elements.start_m[first], axis=1)
                   / (C_M_PER_NS / fastest))

    # prompt orders 0 and 1 with the predicted >=2 in the selection
    safety = float(config.predictor_safety)
    stage = perf_counter()
    prompt = transport_prompt(kernel, elements, config.prompt, order0=order0,
                              selection_extra_charge=safety * np.abs(predicted)
                              + np.abs(late_charge) + extra,
                              time_origin_ns=origins if local_origin else None, device=device,
                              chords=chords)
    timings["prompt_seconds"] = perf_counter() - stage
    if not np.allclose(prompt.time_origin_ns, origins, rtol=0, atol=1e-10):
        raise ValueError("prompt and RTE time origins differ")

    # exact zero-frequency >=2 where the threshold decision needs it
    stage = perf_counter()
    charge = np.zeros((n, 3))
    charge[:, :2] = prompt.charge_orders_pe
    prompt_charge = charge[:, :2].sum(axis=1)
    computable = (used > 0) | (prompt.charge_pe > 0)
    candidate = (used > 0) & (prompt_charge + safety * np.abs(predicted) + late_charge + extra
                              >= threshold)
    if threshold == 0:
        candidate = used > 0

    def exact(mask):
        ids = np.flatnonzero(mask)
        values = np.zeros(n)
        if len(ids):
            values[ids] = apply_bands(pack, sources, positions[ids], looks[ids], alpha, area[ids],
                                      (low, high), frequencies=1)[0].real
        return values

    level = np.where(used > 0, 1, 0).astype(np.int8)
    exact_charge = exact(candidate)
    ratio = np.abs(exact_charge[candidate]) / np.maximum(np.abs(predicted[candidate]), 1e-300)
    fallback = bool(np.any(np.abs(exact_charge[candidate])
                           > safety * np.abs(predicted[candidate])))
    if fallback:
        candidate = used > 0
        exact_charge = exact(candidate)
    charge[:, 2] = predicted + late_charge
    charge[candidate, 2] = exact_charge[candidate] + late_charge[candidate]
    level[candidate] = 2
    active = computable & (charge.sum(axis=1) + extra >= threshold)
    if threshold == 0:
        active = computable
    if np.any(active & (level == 1)):
        raise RuntimeError("an OM was activated by a predicted >=2 charge")
    if not np.all(prompt.active[active]) and not fallback:
        raise RuntimeError("prompt bins are missing for an active OM")
    timings["zero_exact_seconds"] = perf_counter() - stage

    # full frequencies for the active OMs, then the readout
    stage = perf_counter()
    spectrum = np.zeros((len(omega), n), complex)
    work = np.flatnonzero(active & (used > 0))
    if len(work):
        spectrum[:, work] = apply_bands(full_pack, sources, positions[work], looks[work], alpha,
                                        area[work], (low, high))
    late_spectra = []
    for shift, group in late_sources:
        values = np.zeros((len(omega), n), complex)
        if len(work):
            values[:, work] = apply_bands(full_pack, group, positions[work], looks[work], alpha,
                                          area[work], (low, high))
        late_spectra.append((shift, values))
    timings["full_apply_seconds"] = perf_counter() - stage
    stage = perf_counter()
    if fallback and not np.all(prompt.computed[active]):
        raise ValueError("prompt did not compute every active OM")
    spectrum[0] = charge[:, 2] - late_charge
    components = np.zeros((n, len(edges) - 1, 3))
    if np.any(active):
        if not np.all(prompt.active[active]):
            redo = transport_prompt(kernel, elements, replace(config.prompt, screen=False),
                                    order0=order0, compute_bins_for_all=True,
                                    time_origin_ns=origins if local_origin else None,
                                    device=device, chords=chords)
            components[active, :, :2] = redo.bins_orders_pe[active]
        else:
            components[active, :, :2] = prompt.bins_orders_pe[active]
        ids = np.flatnonzero(active)
        readout = (route.readout(omega, edges) if route is not None else
                   _readout_basis(kernel, omega, edges, device))
        device_spectrum = torch.as_tensor(np.ascontiguousarray(spectrum[:, ids], np.complex64),
                                          device=device)
        device_origins = torch.as_tensor(origins[ids], device=device)
        components[ids, :, 2] = readout.apply(device_spectrum, device_origins,
                                              precision=readout_precision).double().cpu().numpy()
        for shift, values in late_spectra:
            basis_shift = (route.readout(omega, edges - shift) if route is not None else
                           _readout_basis(kernel, omega, edges - shift, device))
            late = torch.as_tensor(np.ascontiguousarray(values[:, ids], np.complex64),
                                   device=device)
            components[ids, :, 2] += basis_shift.apply(
                late, torch.as_tensor(origins[ids] + shift, device=device),
                precision=readout_precision).double().cpu().numpy()
    timings["readout_seconds"] = perf_counter() - stage
    elapsed = perf_counter() - began
    timings.update(elapsed_seconds=elapsed, cpu_seconds=process_time() - cpu)
    metadata = {
        "route": "lighthit.shower", "timings": timings, "prompt_timings": prompt.metadata["timings"],
        "elapsed_seconds": elapsed, "prompt_elapsed_seconds": timings["prompt_seconds"],
        "ge2_basis": asdict(basis), "band_fold": fold.metadata(), "radial_tail": radial_tail,
        "ge2_points": [int(item.points.shape[0]) for item in sources],
        "delayed_light": {"prompt_delay_ns": float(config.prompt_delay_ns),
                          "groups_in_window": len(late_sources),
                          "steps_in_window_groups": int(sum(int(m.sum()) for m, _ in late_groups)),
                          "steps_beyond_window": int(beyond.sum()),
                          "delayed_charge_pe": float(late_charge.sum())},
        "ge2_charge_level": level.tolist(),
        "ge2_predictor": {"points": int(len(collapsed[0].z_m)),
                          "cell_m": float(config.predictor_cell_m), "safety": safety,
                          "candidates": int(candidate.sum()), "reachable": int(len(reach)),
                          "fallback_to_exact": fallback,
                          "exact_over_prediction_max": float(ratio.max(initial=0.0)),
                          "exact_over_prediction_min": (float(ratio.min()) if len(ratio)
                                                        else None)},
        "radial_window_m": [float(low), float(high)],
        "radial_windowed_modules": int(np.count_nonzero(far)),
        "time_origin": ("supplied event clock" if time_origin_ns is not None else
                        "earliest local segment and fastest group flight" if local_origin
                        else "earliest element start and fastest group flight"),
        "selection_extra_charge_pe": float(extra.sum()),
        "source_elements": len(elements),
        "source_coefficient_fraction_dropped":
            1 - float(elements.coefficient0.sum()) / original_coefficient,
        "active_modules": int(active.sum()), "threshold_pe": threshold,
        "cone_model": "spectral", "spectrum_computed_orders": [2],
        "frequency_nodes_computed": len(omega), "frequency_nodes_readout": len(output_omega),
        "prompt_order1_level": prompt.metadata["order1_level"],
        "backend": "lighthit_amd (gfx950)",
    }
    response = HybridTransportResponse(kernel.detector, omega.copy(), edges.copy(), spectrum,
                                       components, charge, origins, active, "shower", metadata)
    if np.any(near):
        # source points inside the tables' radial minimum: the near-field clamp
        # (shower/api.py:536-538)
        from .near_field import add_near_field
        near_pack = (route.near_pack(kernel) if route is not None else
                     _near_pack(tables, kernel, config, device))
        distance, _ = segment_bounds(positions, elements, None, device, chords)
        candidates = np.flatnonzero(distance < near_pack.radial_low * (1 + 1e-12))
        readout = (route.readout(omega, edges) if route is not None else
                   _readout_basis(kernel, omega, edges, device))
        add_near_field(response, kernel, elements, near_pack, readout, label="shower",
                       candidates=candidates, device=device,
                       precision=readout_precision)
    return response


def _near_pack(tables, kernel, config, device):
    """The per-wavelength order >=2 pack on the route's frequencies, on the device
    (the near-field clamp evaluates it wavelength by wavelength)."""
    from .fold import KernelPack, frequency_indices
    from .ge2 import DevicePack
    pack = tables["pack"]
    if not isinstance(pack, KernelPack):
        pack = KernelPack.load(pack)
    indices = frequency_indices(kernel.config.omega_per_ns, config.frequency_nodes)
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, int(kernel.config.source_degree),
                         indices)
    return DevicePack(bundle, device)


def _cached_readout(cache, lock, kernel, omega, edges, device, limit):
    """A ReadoutBasis from ``cache`` (an OrderedDict guarded by ``lock``), keyed on the
    exact bytes of the frequency grid, the readout grid and the bin edges."""
    from .readout import ReadoutBasis
    omega = np.ascontiguousarray(omega, float)
    edges = np.ascontiguousarray(edges, float)
    readout = getattr(kernel.config, "readout_omega_per_ns", None)
    readout = None if readout is None else np.ascontiguousarray(readout, float)
    key = (omega.tobytes(), edges.tobytes(), None if readout is None else readout.tobytes(),
           str(device))
    with lock:
        basis = cache.get(key)
        if basis is not None:
            cache.move_to_end(key)
            return basis
    basis = ReadoutBasis(omega, edges, readout, device=device)
    with lock:
        cache[key] = basis
        cache.move_to_end(key)
        while len(cache) > limit:
            cache.popitem(last=False)
    return basis


_route_free_readouts = OrderedDict()
_route_free_lock = threading.Lock()


def _readout_basis(kernel, omega, edges, device):
    """Readout basis for calls without a route object (small shared LRU cache)."""
    return _cached_readout(_route_free_readouts, _route_free_lock, kernel, omega, edges, device,
                           ShowerTransport.READOUT_CACHE)
