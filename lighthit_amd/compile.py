"""Order >=2 source on the GPU: the axial-slice compile of the node bands.

Host mirror of ``compile_axial_slices`` (``_shower_basis.py:188-645``) with the
options of the shower route (``shower/ge2.py:49-55``; stencils centred on the
axis, no point selection) and of ``node_sources`` (``shower/ge2.py:35-67``).
The O(N) scalars that fix the lattice (chord node rule, emission-time
reference, slice origin, ticks per slice, slice range) are set here exactly as
the reference sets them; the lattice, the geometry decision, the deposit, the
phase sums and the stencil run in ``lh_compile_layout`` and
``lh_compile_channels``. The result stays on the device
(:class:`CompiledBand`); :func:`compile_axial_slices` returns the reference's
numpy ``AxialSource`` for callers that want host arrays.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .axial import AxialSource, AxisFrame, channel_layout, dot3


class _Field(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "start_m", "direction", "length_m", "start_ns", "end_ns", "coefficient0",
        "coefficient2", "beta", "photons", "cone_cosine", "mask")] + [
        ("scale0", ctypes.c_double), ("scale2", ctypes.c_double),
        ("phase_index", ctypes.c_double), ("count", ctypes.c_int)]


class _Lattice(ctypes.Structure):
    _fields_ = [("centre", ctypes.c_double * 3), ("axis", ctypes.c_double * 3),
                ("first", ctypes.c_double * 3), ("second", ctypes.c_double * 3)] + [
        (name, ctypes.c_double) for name in ("z_origin_m", "reference_ns", "cell_m",
                                             "time_step_ns", "stencil_m",
                                             "stencil_threshold_m")] + [
        (name, ctypes.c_int) for name in ("degree", "azimuthal_degree", "axial_order",
                                          "moments", "element_order", "time_deposit", "kmax",
                                          "slice0", "positions")]


_P = ctypes.c_void_p
_lib.SIGNATURES["lh_compile_workspace_bytes"] = (ctypes.c_size_t,
                                                 [ctypes.POINTER(_Lattice), ctypes.c_int])
_lib.SIGNATURES["lh_compile_layout"] = (ctypes.c_int, [ctypes.POINTER(_Field),
                                                       ctypes.POINTER(_Lattice), _P, _P, _P,
                                                       ctypes.c_size_t, _P])
_lib.SIGNATURES["lh_compile_channels"] = (ctypes.c_int, [
    ctypes.POINTER(_Field), ctypes.POINTER(_Lattice), _P, ctypes.c_int, ctypes.c_double,
    ctypes.c_int, _P, _P, _P, ctypes.c_size_t, _P])

INFO_SLICES, INFO_WIDE, INFO_POINTS, INFO_ROWS, INFO_FLAGS = range(5)
FLAG_RANGE = 1

_TIME_DEPOSIT = {"nearest": 0, "linear": 1}


@dataclass(frozen=True)
class BasisOptions:
    """The axial-basis choices of ``ShowerBasis`` (``shower_basis.py:30-75``) that the
    compile reads."""
    cell_m: float = 0.25
    axial_order: int = 3
    transverse: str = "stencil"
    stencil_m: float = 0.2
    stencil_threshold_m: float = 0.03
    time_step_ns: float = 0.02
    time_deposit: str = "linear"

    def __post_init__(self):
        if not (self.cell_m > 0 and self.time_step_ns > 0 and self.stencil_m > 0
                and self.stencil_threshold_m >= 0):
            raise ValueError("cell_m, time_step_ns and stencil_m must be positive")
        if self.axial_order not in (1, 3):
            raise ValueError("axial_order must be 1 or 3")
        if self.transverse not in ("collapse", "stencil"):
            raise ValueError("transverse must be 'collapse' or 'stencil'")
        if self.time_deposit not in _TIME_DEPOSIT:
            raise ValueError("the GPU compile provides the linear and nearest time lattices")


@dataclass
class CompiledBand:
    """One compiled band on the device: points (P, 3) in the frame (first, second,
    axis) and channels (P, C, W) complex64."""
    frame: AxisFrame
    points: object
    channels: object
    degree: int
    azimuthal_degree: int
    reference_ns: float
    summary: dict = field(default_factory=dict)

    def to_points_m(self):
        """Point positions in detector coordinates (host array)."""
        p = self.points.double().cpu().numpy()
        f = self.frame
        return (np.asarray(f.centre_m)[None, :] + p[:, 2:3] * np.asarray(f.axis)[None, :]
                + p[:, 0:1] * np.asarray(f.first)[None, :]
                + p[:, 1:2] * np.asarray(f.second)[None, :])

    def to_source(self):
        """The reference's ``AxialSource`` with host arrays (complex128 channels)."""
        points = self.points.double().cpu().numpy()
        return AxialSource(self.frame, points[:, 2].copy(), points[:, :2].copy(),
                           self.channels.cpu().numpy().astype(np.complex128), self.degree,
                           self.azimuthal_degree, self.reference_ns, dict(self.summary))


class DeviceChords:
    """Chord arrays of one source on the device, shared by every band's compile."""

    def __init__(self, elements, device="cuda"):
        import torch
        self.count = len(elements)
        self.device = device
        self.host = elements

        spectral = hasattr(elements, "coefficient0")
        names = ["start_m", "direction", "length_m", "start_ns", "end_ns"]
        names += (["coefficient0", "coefficient2", "beta"] if spectral
                  else ["photons", "cone_cosine"])
        # one host-to-device copy for every array
        arrays = _lib.upload([np.asarray(getattr(elements, n), float) for n in names], device)
        got = dict(zip(names, arrays))
        self.start_m, self.direction = got["start_m"], got["direction"]
        self.length_m, self.start_ns, self.end_ns = got["length_m"], got["start_ns"], got["end_ns"]
        self.coefficient0 = got.get("coefficient0")
        self.coefficient2 = got.get("coefficient2")
        self.beta = got.get("beta")
        self.photons = got.get("photons")
        self.cone = got.get("cone_cosine")


def _address(tensor):
    return None if tensor is None else ctypes.c_void_p(tensor.data_ptr())


def _workspace(device, nbytes, stream=None):
    return _lib.workspace("compile", device, nbytes, stream)


def _lattice(chords_host, taken, frame, degree, azimuthal_degree, omega, options,
             element_order):
    """Scalars of the lattice, as ``_shower_basis.py:285-311`` sets them."""
    length = np.asarray(chords_host.length_m, float)[taken]
    if element_order == "auto":
        element_order = 1 if float(np.max(length)) <= options.cell_m / 10 else 2
    element_order = int(element_order)
    if element_order not in (1, 2):
        raise ValueError("element_order must be 1, 2 or 'auto' on the GPU")
    nodes, _ = np.polynomial.legendre.leggauss(element_order)
    nodes = (nodes + 1) / 2
    start_ns = np.asarray(chords_host.start_ns, float)[taken]
    end_ns = np.asarray(chords_host.end_ns, float)[taken]
    start = np.asarray(chords_host.start_m, float)[taken]
    direction = np.asarray(chords_host.direction, float)[taken]
    reference = float(start_ns.min())
    charge_only = len(omega) == 1 and omega[0] == 0
    margin = 1 if options.axial_order == 1 else 2
    centre, axis = np.asarray(frame.centre_m, float), np.asarray(frame.axis, float)

    def node_z(fraction):
        points = start + (fraction * length)[:, None] * direction
        return dot3(points - centre, axis)
    z_origin = float(node_z(0.5).min()) - margin * options.cell_m
    times = [start_ns + fraction * (end_ns - start_ns) for fraction in nodes]
    if min(float(t.min()) for t in times) < reference:
        raise ValueError("chord end times precede their start times")
    time_max = max(float(t.max()) for t in times)
    kmax = 1 if charge_only else int(np.floor((time_max - reference) / options.time_step_ns)) + 3
    zb = np.concatenate([np.floor((node_z(f) - z_origin) / options.cell_m) for f in nodes])
    low, high = (0, 1) if options.axial_order == 1 else (-1, 2)
    slice0 = int(zb.min()) + low - 1
    positions = int(zb.max()) + high - slice0 + 2
    lattice = _Lattice()
    for name in ("centre", "axis", "first", "second"):
        vector = np.asarray(getattr(frame, name if name != "centre" else "centre_m"), float)
        getattr(lattice, name)[:] = [float(v) for v in vector]
    lattice.z_origin_m, lattice.reference_ns = z_origin, reference
    lattice.cell_m, lattice.time_step_ns = float(options.cell_m), float(options.time_step_ns)
    lattice.stencil_m = float(options.stencil_m)
    lattice.stencil_threshold_m = float(options.stencil_threshold_m)
    lattice.degree, lattice.azimuthal_degree = int(degree), int(azimuthal_degree)
    lattice.axial_order = int(options.axial_order)
    lattice.moments = 1 if options.transverse == "collapse" else 5
    lattice.element_order = element_order
    lattice.time_deposit = 2 if charge_only else _TIME_DEPOSIT[options.time_deposit]
    lattice.kmax, lattice.slice0, lattice.positions = kmax, slice0, positions
    return lattice


def compile_band(chords, frame, degree, omega, *, azimuthal_degree=4, options=None,
                 wavelength_nm=None, phase_index=None, mask=None, scale=1.0,
                 element_order="auto", omega_device=None, stream=None):
    """Compile one band of ``chords`` (a :class:`DeviceChords`) on the device.

    Spectral chords take the band's ``wavelength_nm`` and ``phase_index``
    (photons ``C0/lambda^2 + C2/(lambda^2 n^2)``, cone ``1/(beta n)``, elements
    with ``beta n <= 1`` left out, ``shower/ge2.py:21-32``); plain fields carry
    their photons and cones. ``mask`` (bool per element) restricts the
    elements; ``scale`` multiplies the channels (the fold's band weight).
    Returns a :class:`CompiledBand`, or None when no element emits.
    """
    import torch
    options = options or BasisOptions()
    omega = np.atleast_1d(np.asarray(omega, float))
    host = chords.host
    taken = np.ones(chords.count, bool) if mask is None else np.asarray(mask, bool).copy()
    spectral = chords.photons is None
    if spectral:
        if wavelength_nm is None or phase_index is None:
            raise ValueError("spectral chords need the band's wavelength and phase index")
        taken &= np.asarray(host.beta, float) * float(phase_index) > 1
    if not np.any(taken):
        return None
    # the lattice depends on the taken elements, the frame and the options only:
    # bands with the same elements share it
    key = (taken.tobytes(), id(frame), int(degree), int(azimuthal_degree),
           bool(len(omega) == 1 and omega[0] == 0), options, str(element_order))
    cache = chords.__dict__.setdefault("_lattices", {})
    entry = cache.get(key)
    if entry is None or entry[0] is not frame:
        entry = cache[key] = (frame, _lattice(host, taken, frame, degree, azimuthal_degree,
                                              omega, options, element_order))
    lattice = entry[1]
    device = chords.device
    mask_dev = (None if mask is None else
                torch.as_tensor(np.asarray(mask, np.uint8), device=device))
    descriptor = _Field()
    for name in ("start_m", "direction", "length_m", "start_ns", "end_ns", "coefficient0",
                 "coefficient2", "beta", "photons"):
        setattr(descriptor, name, _address(getattr(chords, name)))
    descriptor.cone_cosine = _address(chords.cone)
    descriptor.mask = _address(mask_dev)
    if spectral:
        scale0 = 1.0 / float(wavelength_nm) ** 2
        descriptor.scale0 = scale0
        descriptor.scale2 = scale0 / float(phase_index) ** 2
        descriptor.phase_index = float(phase_index)
    descriptor.count = chords.count
    lib = _lib.load()
    fn = lib.lh_compile_workspace_bytes
    fn.restype, fn.argtypes = _lib.SIGNATURES["lh_compile_workspace_bytes"]
    need = int(fn(ctypes.byref(lattice), chords.count))
    if need == 0:
        raise ValueError(lib.lh_last_error().decode(errors="replace"))
    ws = _workspace(device, need, stream)
    info = torch.zeros(8, dtype=torch.int32, device=device)
    points = torch.empty((5 * lattice.positions, 3), dtype=torch.float64, device=device)
    handle = _lib.stream_handle(stream)
    with _lib.stage("compile_layout"):
        _lib.call("lh_compile_layout", ctypes.byref(descriptor), ctypes.byref(lattice),
                  _address(info), _address(points), _address(ws), ctypes.c_size_t(ws.numel()),
                  handle)
    counts = info.cpu().numpy()
    if counts[INFO_FLAGS] & FLAG_RANGE:
        raise RuntimeError("a chord node fell outside the slice lattice")
    npoints, rows = int(counts[INFO_POINTS]), int(counts[INFO_ROWS])
    nch = int(channel_layout(int(degree), min(int(degree), int(azimuthal_degree)))[-1])
    deposit = torch.empty(rows * (lattice.moments * nch + 2 * len(omega)), dtype=torch.float64,
                          device=device)
    channels = torch.empty((npoints, nch, len(omega)), dtype=torch.complex64, device=device)
    if omega_device is None:
        omega_device = torch.as_tensor(omega, device=device)
    with _lib.stage("compile_channels"):
        _lib.call("lh_compile_channels", ctypes.byref(descriptor), ctypes.byref(lattice),
                  _address(omega_device), len(omega), float(scale), rows, _address(deposit),
                  _address(channels), _address(ws), ctypes.c_size_t(ws.numel()), handle)
    summary = dict(elements=int(taken.sum()), slices=int(counts[INFO_SLICES]),
                   stencil_slices=int(counts[INFO_WIDE]), time_rows=rows,
                   element_order=lattice.element_order, z_origin_m=lattice.z_origin_m,
                   kmax=lattice.kmax)
    return CompiledBand(frame, points[:npoints], channels, int(degree), int(azimuthal_degree),
                        lattice.reference_ns, summary)


def compile_axial_slices(field_, degree, omega_per_ns, *, azimuthal_degree=4, cell_m=0.25,
                         axial_order=3, transverse="stencil", stencil_m=0.2,
                         stencil_threshold_m=0.0, time_step_ns=0.02, time_deposit="linear",
                         element_order="auto", frame=None, device="cuda"):
    """Drop-in for ``compile_axial_slices`` (``_shower_basis.py:188``) for one field;
    returns an ``AxialSource`` with host arrays."""
    options = BasisOptions(cell_m, axial_order, transverse, stencil_m, stencil_threshold_m,
                           time_step_ns, time_deposit)
    frame = frame or AxisFrame.of(field_)
    chords = DeviceChords(field_, device)
    band = compile_band(chords, frame, degree, omega_per_ns, azimuthal_degree=azimuthal_degree,
                        options=options, element_order=element_order)
    if band is None:
        raise ValueError("no element emits")
    return band.to_source()
