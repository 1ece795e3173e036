"""Order 1 (single scattering) on the GPU: grouped pencils with fused bands.

Host mirror of the spectral-cone order-1 block of ``transport_prompt``
(``prompt.py:670-851``) on the grouped path (``_order1``,
``prompt.py:1020-1258``): evaluation units (``muon/secondary_first.py:184-204``),
the coarse screen, the fine pass with bins and the widening step. Pencil
compression and the grouped evaluation run in ``lh_order1_*``; only the unit
masks, the lattice scalars and the threshold decisions are host code.
"""
from __future__ import annotations

import ctypes
import math
import threading
from dataclasses import dataclass

import numpy as np

from . import _lib

C_M_PER_NS = 0.299792458
_P = ctypes.c_void_p


class _Chords(ctypes.Structure):
    _fields_ = [(n, _P) for n in ("start_m", "direction", "length_m", "coefficient0",
                                  "coefficient2", "cone_cosine", "start_ns", "end_ns", "index",
                                  "sub_offset")] + [("count", ctypes.c_int)]


class _Grid(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_double * 3), ("dims", ctypes.c_longlong * 3),
                ("cell_m", ctypes.c_double), ("pixel_face", ctypes.c_int),
                ("ring_nodes", ctypes.c_int), ("parent_face", ctypes.c_int),
                ("subsegments", ctypes.c_longlong)]


class _Pencils(ctypes.Structure):
    _fields_ = [(n, _P) for n in ("position", "direction", "w0", "w2", "time", "major",
                                  "var_major", "var_minor", "pos_cov", "parent", "next")] + [
        ("count", ctypes.c_int)]


class _Oms(ctypes.Structure):
    _fields_ = [(n, _P) for n in ("positions", "looks", "areas", "origins", "work")] + [
        ("work_count", ctypes.c_int)]


class _Optics(ctypes.Structure):
    _fields_ = [(n, _P) for n in ("mus", "mut", "speed", "s0w", "s2w", "edges", "eedge", "itab",
                                  "coef", "ipow", "jpow")] + [
        (n, ctypes.c_int) for n in ("bands", "bins", "n_mono", "max_power", "direct_bins",
                                    "max_path_nodes")] + [
        (n, ctypes.c_double) for n in ("g", "y_g", "path_step", "path_step_tail",
                                       "tail_exponent", "smoothing", "mut_min", "speed_max",
                                       "bin_inverse_width")]


class _Merge(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("theta_lo", "theta_ratio", "tau_step",
                                               "log_r_step", "speed_ref")]


_S = ctypes.c_size_t
_I = ctypes.c_int
_lib.SIGNATURES.update({
    "lh_order1_set_tables": (_I, [_P, _P, _P]),
    "lh_order1_compress_workspace_bytes": (_S, [ctypes.POINTER(_Chords), ctypes.POINTER(_Grid)]),
    "lh_order1_compress_keys": (_I, [ctypes.POINTER(_Chords), ctypes.POINTER(_Grid), _P, _P, _S,
                                     _P]),
    "lh_order1_compress_pencils": (_I, [ctypes.POINTER(_Chords), ctypes.POINTER(_Grid), _I,
                                        ctypes.POINTER(_Pencils), _P, _P, _S, _P]),
    "lh_order1_grouped_workspace_bytes": (_S, [ctypes.POINTER(_Pencils), ctypes.POINTER(_Oms),
                                               ctypes.POINTER(_Optics)]),
    "lh_order1_group_keys": (_I, [ctypes.POINTER(_Pencils), ctypes.POINTER(_Oms),
                                  ctypes.POINTER(_Optics), ctypes.POINTER(_Merge), _P, _P, _S,
                                  _P]),
    "lh_order1_group_scratch_bytes": (_S, [ctypes.POINTER(_Pencils), ctypes.POINTER(_Oms),
                                           ctypes.POINTER(_Optics), _I, ctypes.c_longlong]),
    "lh_order1_group_charge": (_I, [ctypes.POINTER(_Pencils), ctypes.POINTER(_Oms),
                                    ctypes.POINTER(_Optics), ctypes.POINTER(_Merge), _I, _I, _P,
                                    _S, _I, _P, _P, _P, _P, _S, _P]),
})

_GL48 = np.polynomial.legendre.leggauss(48)
_TAU = 0.25 * math.pi * (_GL48[0] + 1.0)
_TANGENT = np.tan(_TAU)
_WEIGHTED = 0.25 * math.pi * _GL48[1] * (1.0 + _TANGENT * _TANGENT)
_TAYLOR = np.asarray([[1.0 / (math.factorial(n) * (n + j + 1.0)) for n in range(24)]
                      for j in range(3)], dtype=np.float64)
_tables_ready = set()


_tables_lock = threading.Lock()


def _ensure_tables(device):
    import torch
    key = str(torch.device(device))
    if key in _tables_ready:
        return
    with _tables_lock:
        if key in _tables_ready:
            return
        with torch.cuda.device(torch.device(device)):
            _lib.call("lh_order1_set_tables", _TAYLOR.ctypes.data_as(_P),
                      _TANGENT.ctypes.data_as(_P), _WEIGHTED.ctypes.data_as(_P))
            torch.cuda.synchronize(torch.device(device))
        _tables_ready.add(key)


def _size(name, *args):
    lib = _lib.load()
    fn = getattr(lib, name)
    fn.restype, fn.argtypes = _lib.SIGNATURES[name]
    value = int(fn(*args))
    if value == 0:
        raise ValueError(lib.lh_last_error().decode(errors="replace"))
    return value


def _addr(tensor):
    return None if tensor is None else ctypes.c_void_p(tensor.data_ptr())


def _workspace(tag, device, nbytes, stream=None):
    return _lib.workspace("order1_" + tag, device, nbytes, stream)


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
    fused_pixel_face: int = 96
    fused_ring_nodes: int = 192
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
    """Monomials of ``A(x) = sum alpha_l (2l+1) P_l(x) / (4 pi)`` (``prompt.py:380-399``)."""
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
    return np.asarray(coef, float), np.asarray(ipow), np.asarray(jpow), int(degree)


def spectral_tables(medium, detector, quadrature):
    """Per-wavelength optics and spectral weights (``prompt.py:412-434``)."""
    wavelengths = np.asarray(quadrature.wavelength_nm, float)
    weights = np.asarray(quadrature.weight_nm, float)
    sample = medium.sample(wavelengths)
    efficiency = np.array([float(detector.spectral_weight(float(w))) for w in wavelengths])
    mus = np.asarray(sample["scattering_per_m"], float)
    mua = np.asarray(sample["absorption_per_m"], float)
    phase = np.asarray(sample["phase_index"], float)
    return {"wavelength_nm": wavelengths, "weight_nm": weights, "efficiency": efficiency,
            "mus": mus, "mua": mua, "mut": mua + mus,
            "speed": C_M_PER_NS / np.asarray(sample["group_index"], float), "phase_index": phase,
            "group_index": np.asarray(sample["group_index"], float),
            "s0w": weights * efficiency / wavelengths ** 2,
            "s2w": weights * efficiency / (wavelengths ** 2 * phase ** 2)}


def _ifuncs(z):
    """``(i0, i1, i2)`` of every entry of ``z``: the integrals of ``t^k e^{z t}``
    over [0, 1], by their Taylor series inside |z| < 1 and in closed form
    outside (``_prompt_numba.py:160-183``)."""
    z = np.asarray(z, float)
    out = np.empty(z.shape + (3,))
    small = np.abs(z) < 1.0
    zs = z[small]
    for j in range(3):
        acc = np.full(zs.shape, _TAYLOR[j, 23])
        for n in range(22, -1, -1):
            acc = acc * zs + _TAYLOR[j, n]
        out[small, j] = acc
    zl = z[~small]
    ez = np.exp(zl)
    out[~small, 0] = (ez - 1.0) / zl
    out[~small, 1] = (ez * (zl - 1.0) + 1.0) / (zl * zl)
    out[~small, 2] = (ez * (zl * zl - 2.0 * zl + 2.0) - 2.0) / (zl * zl * zl)
    return out


_optics_cache = {}


def device_optics(tables, alpha, g, edges, settings, device="cuda"):
    """The :class:`DeviceOptics` of these inputs, built once per process and
    kept resident (its tables depend on the medium, detector and bins only)."""
    arrays = [np.asarray(tables[k], float).tobytes() for k in ("mus", "mut", "speed", "s0w",
                                                               "s2w")]
    key = (*arrays, np.asarray(alpha, float).tobytes(), float(g),
           np.asarray(edges, float).tobytes(), int(settings.direct_bins),
           int(settings.max_path_nodes), float(settings.path_step),
           float(settings.path_step_tail), float(settings.tail_exponent),
           float(settings.smoothing_factor), str(device))
    optics = _optics_cache.get(key)
    if optics is None:
        if len(_optics_cache) >= 32:
            _optics_cache.clear()
        optics = _optics_cache[key] = DeviceOptics(tables, alpha, g, edges, settings, device)
    return optics


class DeviceOptics:
    """Band optics, bin edges and acceptance monomials of one evaluation unit."""

    def __init__(self, tables, alpha, g, edges, settings, device="cuda"):
        import torch
        self.tables = tables
        edges = np.asarray(edges, float)
        kappa = -tables["mut"] * tables["speed"]
        eedge = np.exp(np.outer(kappa, edges))
        widths = np.diff(edges)
        itab = _ifuncs(np.outer(kappa, widths))
        coef, ipow, jpow, degree = acceptance_monomials(alpha)

        def up(a, dtype=torch.float64):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)
        self._keep = dict(mus=up(tables["mus"]), mut=up(tables["mut"]), speed=up(tables["speed"]),
                          s0w=up(tables["s0w"]), s2w=up(tables["s2w"]), edges=up(edges),
                          eedge=up(eedge), itab=up(itab), coef=up(coef),
                          ipow=up(ipow, torch.int32), jpow=up(jpow, torch.int32))
        s = _Optics()
        for name, tensor in self._keep.items():
            setattr(s, name, _addr(tensor))
        s.bands, s.bins, s.n_mono, s.max_power = len(kappa), len(widths), len(coef), degree
        s.direct_bins, s.max_path_nodes = int(settings.direct_bins), int(settings.max_path_nodes)
        s.g = float(g)
        s.y_g = (1.0 - abs(g)) / (1.0 + abs(g)) if g != 0 else 1.0
        s.path_step, s.path_step_tail = float(settings.path_step), float(settings.path_step_tail)
        s.tail_exponent, s.smoothing = float(settings.tail_exponent), float(settings.smoothing_factor)
        s.mut_min, s.speed_max = float(np.min(tables["mut"])), float(np.max(tables["speed"]))
        widths = np.diff(edges)
        equal = bool(np.all(np.abs(widths - widths[0]) <= 1e-12 * abs(widths[0])))
        s.bin_inverse_width = 1.0 / float(widths[0]) if equal and widths[0] > 0 else 0.0
        self.struct = s
        self.n_mono = len(coef)
        self.bins = len(widths)
        self.speed_ref = float(np.max(tables["speed"]))


class DeviceOMs:
    """OM positions, looks, areas and time origins on the device (all OMs)."""

    def __init__(self, positions, looks, areas, origins, device="cuda"):
        import torch

        self.count = len(positions)
        self.device = device
        self._keep = _lib.upload([positions, looks, areas, origins], device)

    def struct(self, work):
        import torch
        work_dev = torch.as_tensor(np.ascontiguousarray(work, np.int32), device=self.device)
        s = _Oms(*[_addr(t) for t in self._keep], _addr(work_dev), len(work))
        return s, work_dev


class DevicePencils:
    """Pencils of one unit on the device (screen pencils carry their hierarchy)."""

    def __init__(self, count, device, screen):
        import torch
        shapes = dict(position=3, direction=3, w0=1, w2=1, time=1, major=3, var_major=1,
                      var_minor=1, pos_cov=6)
        self.arrays = {k: torch.empty((count, n) if n > 1 else (count,), dtype=torch.float64,
                                      device=device) for k, n in shapes.items()}
        self.parent = torch.empty(count, dtype=torch.int64, device=device) if screen else None
        self.next = torch.empty(count, dtype=torch.int64, device=device) if screen else None
        self.capacity = count
        self.count = count

    def struct(self):
        a = self.arrays
        return _Pencils(*[_addr(a[k]) for k in ("position", "direction", "w0", "w2", "time",
                                                "major", "var_major", "var_minor", "pos_cov")],
                        _addr(self.parent), _addr(self.next), int(self.count))

    def host(self):
        out = {k: v[:self.count].cpu().numpy() for k, v in self.arrays.items()}
        if self.parent is not None:
            out["hierarchy_parent"] = self.parent[:self.count].cpu().numpy()
            out["next_node"] = self.next[:self.count].cpu().numpy()
        return out


class DeviceUnitChords:
    """Chords of one evaluation unit (a subset of the source) with its frozen cone."""

    def __init__(self, chords, mask, cone_phase):
        import torch
        host = chords.host
        self.chords = chords
        self.mask = np.asarray(mask, bool)
        self.index = np.flatnonzero(self.mask).astype(np.int32)
        beta = np.asarray(host.beta, float)
        cone = 1.0 / (beta * float(cone_phase))
        device = chords.device
        self.cone = torch.as_tensor(cone, device=device)
        self.index_dev = torch.as_tensor(self.index, device=device)
        self.length = np.asarray(host.length_m, float)[self.index]
        self.start = np.asarray(host.start_m, float)[self.index]
        self.direction = np.asarray(host.direction, float)[self.index]
        self.count = len(self.index)

    def struct(self, cell_m):
        import torch
        a_step = 0.5 * cell_m
        n_a = np.maximum(1, np.ceil(self.length / a_step)).astype(np.int64)
        offsets = np.r_[0, np.cumsum(n_a)].astype(np.int64)
        self._offsets = torch.as_tensor(offsets, device=self.chords.device)
        c = self.chords
        s = _Chords(_addr(c.start_m), _addr(c.direction), _addr(c.length_m),
                    _addr(c.coefficient0), _addr(c.coefficient2), _addr(self.cone),
                    _addr(c.start_ns), _addr(c.end_ns), _addr(self.index_dev),
                    _addr(self._offsets), self.count)
        return s, int(offsets[-1])

    def grid(self, cell_m, pixel_face, ring_nodes, parent_face=0):
        if getattr(self, "_box", None) is None:
            from .axial import column_bounds
            ends = self.start + self.length[:, None] * self.direction
            (s_low, s_high), (e_low, e_high) = column_bounds(self.start), column_bounds(ends)
            self._box = (np.minimum(s_low, e_low), np.maximum(s_high, e_high))
        low = self._box[0] - cell_m
        high = self._box[1] + cell_m
        dims = np.ceil((high - low) / cell_m).astype(np.int64) + 1
        if float(np.prod(dims.astype(float))) * 6 * int(pixel_face) ** 2 >= 2.0 ** 62:
            raise ValueError("cell grid too large for 64-bit pencil keys; increase cell_m")
        g = _Grid()
        g.origin[:] = [float(v) for v in low]
        g.dims[:] = [int(v) for v in dims]
        g.cell_m, g.pixel_face, g.ring_nodes = float(cell_m), int(pixel_face), int(ring_nodes)
        g.parent_face = int(parent_face)
        return g


def compress(unit, cell_m, pixel_face, ring_nodes, *, parent_face=0, stream=None):
    """Pencils of ``unit`` (``compress`` / ``compress_screen``) on the device."""
    import torch
    device = unit.chords.device
    chords, subsegments = unit.struct(cell_m)
    grid = unit.grid(cell_m, pixel_face, ring_nodes, parent_face)
    grid.subsegments = subsegments
    need = _size("lh_order1_compress_workspace_bytes", ctypes.byref(chords), ctypes.byref(grid))
    ws = _workspace("compress", device, need, stream)
    info = torch.zeros(8, dtype=torch.int32, device=device)
    handle = _lib.stream_handle(stream)
    with _lib.stage("order1_compress"):
        _lib.call("lh_order1_compress_keys", ctypes.byref(chords), ctypes.byref(grid),
                  _addr(info), _addr(ws), ctypes.c_size_t(ws.numel()), handle)
    runs = int(info[0].item())
    pencils = DevicePencils(2 * runs if parent_face else runs, device, bool(parent_face))
    out = pencils.struct()
    _lib.call("lh_order1_compress_pencils", ctypes.byref(chords), ctypes.byref(grid), runs,
              ctypes.byref(out), _addr(info), _addr(ws), ctypes.c_size_t(ws.numel()), handle)
    pencils.count = runs + (int(info[1].item()) if parent_face else 0)
    return pencils


_work_counts = {}


def count_work(device="cuda", on=True):
    """Start (or stop) counting the grouped passes' work on ``device``: profile
    terms, time panels, bin events and group members (for cost models; off by
    default)."""
    import torch
    key = str(torch.device(device))
    if on:
        _work_counts[key] = torch.zeros(4, dtype=torch.int64, device=device)
    else:
        _work_counts.pop(key, None)


def work_counts(device="cuda"):
    """``(profile_terms, panels, bin_events)`` counted since :func:`count_work`."""
    import torch
    counts = _work_counts.get(str(torch.device(device)))
    return None if counts is None else tuple(int(v) for v in counts.cpu().tolist())


#: (pencil, OM) entries of one grouped call: its sort takes ~40 bytes per entry
MAX_GROUPED_ENTRIES = 1 << 28
#: OMs of one grouped call (the library's limit)
MAX_GROUPED_OMS = 65535


def grouped(pencils, oms, work, optics, merge, *, compute_bins=True, charge=None, bins=None,
            stream=None, pool_nodes=-1):
    """Grouped (or screen) order 1 of OMs ``work``: fills ``charge`` (and ``bins``)
    device tensors at those OMs. ``pool_nodes`` sizes the path-node pool of the
    wide batch kernels (-1: the library's default; 0: the single-kernel form)."""
    import torch
    device = oms.device
    if charge is None:
        charge = torch.zeros(oms.count, dtype=torch.float64, device=device)
    if compute_bins and bins is None:
        bins = torch.zeros((oms.count, optics.bins), dtype=torch.float64, device=device)
    if not len(work):
        return charge, bins
    per_om = (2 if pencils.parent is not None else 1) * max(int(pencils.count), 1)
    most = max(1, min(MAX_GROUPED_OMS, MAX_GROUPED_ENTRIES // per_om))
    if len(work) > most:
        # OMs are independent: a long list runs in parts of bounded size
        for first in range(0, len(work), most):
            grouped(pencils, oms, work[first:first + most], optics, merge,
                    compute_bins=compute_bins, charge=charge, bins=bins, stream=stream,
                    pool_nodes=pool_nodes)
        return charge, bins
    _ensure_tables(device)
    pen = pencils.struct()
    om, work_dev = oms.struct(work)
    m = _Merge(*[float(v) for v in merge], optics.speed_ref)
    need = _size("lh_order1_grouped_workspace_bytes", ctypes.byref(pen), ctypes.byref(om),
                 ctypes.byref(optics.struct))
    ws = _workspace("grouped", device, need, stream)
    info = torch.zeros(8, dtype=torch.int32, device=device)
    handle = _lib.stream_handle(stream)
    name = "order1_screen" if pencils.parent is not None else "order1_grouped"
    with _lib.stage(name + "_keys"):
        _lib.call("lh_order1_group_keys", ctypes.byref(pen), ctypes.byref(om),
                  ctypes.byref(optics.struct), ctypes.byref(m), _addr(info), _addr(ws),
                  ctypes.c_size_t(ws.numel()), handle)
    groups, key_bits = (int(v) for v in info[:2].tolist())
    scratch_bytes = _size("lh_order1_group_scratch_bytes", ctypes.byref(pen), ctypes.byref(om),
                          ctypes.byref(optics.struct), groups, int(pool_nodes))
    scratch = _workspace("group_scratch", device, scratch_bytes, stream)
    with _lib.stage(name + "_charge"):
        _lib.call("lh_order1_group_charge", ctypes.byref(pen), ctypes.byref(om),
                  ctypes.byref(optics.struct), ctypes.byref(m), groups, key_bits, _addr(scratch),
                  ctypes.c_size_t(scratch_bytes), 1 if compute_bins else 0, _addr(charge),
                  _addr(bins) if compute_bins else None,
                  _addr(_work_counts.get(str(torch.device(device)))),
                  _addr(ws), ctypes.c_size_t(ws.numel()), handle)
    del work_dev
    return charge, bins


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


def spectral_order1(chords, oms, work, tables, alpha, g, edges, threshold, order0_charge,
                    selection_extra_charge, settings=None):
    """Order 1 of a spectral-cone element source (``prompt.py:670-851``) on the GPU:
    ``(charge (N,), bins (N, nb), level (N,))`` as numpy arrays.

    ``chords`` is a :class:`lighthit_amd.compile.DeviceChords` of the valid
    elements, ``oms`` a :class:`DeviceOMs`."""
    import torch
    settings = settings or PromptSettings()
    n = oms.count
    nb = len(edges) - 1
    work = np.asarray(work, np.int64)
    device = oms.device
    beta = np.asarray(chords.host.beta, float)
    band_valid = [beta * phase > 1 for phase in tables["phase_index"]]
    tasks = []
    for unit in evaluation_units(tables, band_valid, True):
        if not np.any(unit["valid"]):
            continue
        optics = device_optics(unit["tables"], alpha, g, edges, settings, device)
        face = settings.fused_pixel_face if unit["fused"] else settings.pixel_face
        ring = settings.fused_ring_nodes if unit["fused"] else settings.ring_nodes
        tasks.append((DeviceUnitChords(chords, unit["valid"], unit["cone"]), optics, face, ring))
    extra = (np.zeros(n) if selection_extra_charge is None
             else np.asarray(selection_extra_charge, float))
    level = np.zeros(n, np.int8)
    charge_dev = torch.zeros(n, dtype=torch.float64, device=device)
    bins_dev = torch.zeros((n, nb), dtype=torch.float64, device=device)
    global_screen = selection_extra_charge is not None and settings.screen and threshold > 0
    coarse_merge = (settings.merge_theta_min, 0.3, 2.0, settings.merge_log_r)
    fine_merge = (settings.merge_theta_min, settings.merge_theta_ratio, settings.merge_tau_ns,
                  settings.merge_log_r)
    fine_work = work
    coarse_total = None
    if global_screen and len(work):
        for unit, optics, _, _ in tasks:
            coarse = compress(unit, settings.screen_cell_m,
                              settings.screen_pixel_face * settings.screen_refine_factor,
                              settings.screen_ring_nodes, parent_face=settings.screen_pixel_face)
            part, _ = grouped(coarse, oms, work, optics, coarse_merge, compute_bins=False)
            charge_dev += part
        level[work] = 1
        coarse_total = charge_dev.cpu().numpy()
        fine_work = work[(order0_charge[work] + extra[work]
                          + settings.screen_margin * np.maximum(coarse_total[work], 0.0))
                         >= threshold]
        charge_dev[torch.as_tensor(fine_work, device=device)] = 0.0
    pencil_sets = []
    for unit, optics, face, ring in tasks:
        if not len(fine_work):
            break
        pencils = compress(unit, settings.cell_m, face, ring)
        pencil_sets.append((pencils, optics))
        part, part_bins = grouped(pencils, oms, fine_work, optics, fine_merge)
        index = torch.as_tensor(fine_work, device=device)
        charge_dev[index] += part[index]
        bins_dev += part_bins
    level[fine_work] = 2
    charge = charge_dev.cpu().numpy()
    if global_screen and len(fine_work):
        ratio = coarse_total[fine_work] / np.maximum(charge[fine_work], 1e-300)
        worst = float(np.min(ratio))
        if worst * settings.screen_margin < 1.0:
            wider = 1.5 / max(worst, 1e-3)
            extra_work = work[(level[work] == 1)
                              & (order0_charge[work] + extra[work]
                                 + wider * np.maximum(coarse_total[work], 0.0) >= threshold)]
            if len(extra_work):
                index = torch.as_tensor(extra_work, device=device)
                charge_dev[index] = 0.0
                for pencils, optics in pencil_sets:
                    part, part_bins = grouped(pencils, oms, extra_work, optics, fine_merge)
                    charge_dev[index] += part[index]
                    bins_dev[index] += part_bins[index]
                level[extra_work] = 2
                charge = charge_dev.cpu().numpy()
    # only the OMs of the fine pass carry bins
    bins = np.zeros((n, nb))
    rows = np.flatnonzero(level == 2)
    if len(rows):
        bins[rows] = bins_dev[torch.as_tensor(rows, device=device)].cpu().numpy()
    return charge, bins, level
