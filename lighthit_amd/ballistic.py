"""Stage (a): order-0 direct light to the first glass entry, on the GPU.

Host mirror of the first-entry direct light (``first_entry_ballistic``,
``ballistic_interval.py:268-287``; ``PreparedBallistic.charge``,
``_surface_ballistic.py:91-98``). The optical panels are prepared once per
(medium, detector) (:func:`lighthit_amd.optics.prepare_optics`) and kept on
the device with the acceptance coefficients (:class:`DeviceOptics`); one event
is one launch of ``lh_order0_charge`` over all (OM, element) pairs.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .model import glass_radii, look_directions
from .optics import prepare_optics


class _Optics(ctypes.Structure):
    _fields_ = [("panel_edges", ctypes.c_void_p), ("phase_nodes", ctypes.c_void_p),
                ("group_nodes", ctypes.c_void_p), ("extinction_nodes", ctypes.c_void_p),
                ("efficiency", ctypes.c_void_p), ("major", ctypes.c_void_p),
                ("alpha", ctypes.c_void_p), ("panels", ctypes.c_int),
                ("efficiency_columns", ctypes.c_int), ("major_count", ctypes.c_int),
                ("acceptance_count", ctypes.c_int), ("phase_min", ctypes.c_double),
                ("phase_max", ctypes.c_double)]


class _Oms(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("looks", ctypes.c_void_p),
                ("areas", ctypes.c_void_p), ("radii", ctypes.c_void_p), ("count", ctypes.c_int)]


class _Elements(ctypes.Structure):
    _fields_ = [("start_m", ctypes.c_void_p), ("direction", ctypes.c_void_p),
                ("length_m", ctypes.c_void_p), ("coefficient0", ctypes.c_void_p),
                ("coefficient2", ctypes.c_void_p), ("beta", ctypes.c_void_p),
                ("start_ns", ctypes.c_void_p), ("end_ns", ctypes.c_void_p),
                ("count", ctypes.c_int)]


_lib.SIGNATURES["lh_order0_workspace_bytes"] = (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int])
_lib.SIGNATURES["lh_order0_charge"] = (ctypes.c_int, [ctypes.POINTER(_Optics), ctypes.POINTER(_Oms),
                                                      ctypes.POINTER(_Elements), ctypes.c_void_p,
                                                      ctypes.c_void_p, ctypes.c_size_t,
                                                      ctypes.c_void_p])

_lib.SIGNATURES["lh_order0_charge_noalloc_max_bytes"] = (ctypes.c_size_t, [])
_lib.SIGNATURES["lh_order0_charge_noalloc"] = (ctypes.c_int, [ctypes.POINTER(_Optics),
                                                              ctypes.POINTER(_Oms),
                                                              ctypes.POINTER(_Elements),
                                                              ctypes.c_void_p, ctypes.c_void_p,
                                                              ctypes.c_void_p])

_lib.SIGNATURES["lh_order0_bins_workspace_bytes"] = (ctypes.c_size_t,
                                                     [ctypes.c_int, ctypes.c_int, ctypes.c_int])
_lib.SIGNATURES["lh_order0_bins"] = (ctypes.c_int, [ctypes.POINTER(_Optics), ctypes.POINTER(_Oms),
                                                    ctypes.POINTER(_Elements), ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_size_t,
                                                    ctypes.c_void_p])

_lib.SIGNATURES["lh_order0_count_nodes"] = (ctypes.c_int, [ctypes.c_void_p])
_lib.SIGNATURES["lh_order0_set_rules"] = (ctypes.c_int, [ctypes.c_int, ctypes.c_int])
_lib.SIGNATURES["lh_order0_batch_workspace_bytes"] = (ctypes.c_size_t, [
    ctypes.c_int, ctypes.c_int, ctypes.c_longlong])
_lib.SIGNATURES["lh_order0_batch_chunks"] = (ctypes.c_longlong, [
    ctypes.POINTER(ctypes.c_int), ctypes.c_int])
_lib.SIGNATURES["lh_order0_batch_set"] = (ctypes.c_int, [
    ctypes.POINTER(_Oms), ctypes.POINTER(_Elements), ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
    ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p])
_lib.SIGNATURES["lh_order0_charge_batch"] = (ctypes.c_int, [
    ctypes.POINTER(_Optics), ctypes.POINTER(_Oms), ctypes.c_int, ctypes.c_longlong,
    ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p])


def _query(name, *args):
    lib = _lib.load()
    fn = getattr(lib, name)
    fn.restype, fn.argtypes = _lib.SIGNATURES[name]
    return fn(*args)


def _bucket(n, floor):
    """The smallest power of two at least ``n`` and ``floor``."""
    size = int(floor)
    while size < n:
        size *= 2
    return size

FLAG_GLASS = 1
FLAG_UNSUPPORTED = 2

_node_counter = None


def count_nodes(device="cuda", on=True):
    """Start (or stop) counting the (theta, azimuth) quadrature nodes every
    order-0 launch evaluates (for the cost model; off by default)."""
    global _node_counter
    import torch
    if on:
        _node_counter = torch.zeros(1, dtype=torch.int64, device=device)
        _lib.call("lh_order0_count_nodes", ctypes.c_void_p(_node_counter.data_ptr()))
    else:
        _lib.call("lh_order0_count_nodes", ctypes.c_void_p(0))
        _node_counter = None


def node_count():
    """Nodes counted since :func:`count_nodes` (None when not counting)."""
    return None if _node_counter is None else int(_node_counter.item())


def set_rules(s_nodes=0, azimuth_nodes=0):
    """Quadrature rules of every later order-0 launch. ``s_nodes`` 0 (default)
    picks, for the charge, 3 Gauss nodes per s piece for elements at most a
    tenth of the glass radius long and 6 otherwise, and 6 for the bins; 2, 3, 4
    or 6 force that rule. ``azimuth_nodes`` 0 (default) picks 4 azimuth nodes
    (per azimuth piece of the bins) when the glass radius is at most 1/20 of the
    OM's distance from the element's line and 8 otherwise; 4 or 8 force that
    rule (to compare rule orders)."""
    _lib.call("lh_order0_set_rules", int(s_nodes), int(azimuth_nodes))


def _dev(array, dtype, device):
    import torch
    return torch.as_tensor(np.ascontiguousarray(array), dtype=dtype, device=device)


class DeviceOptics:
    """Optical panels, medium-node lines and acceptance of a detector, on the device."""

    def __init__(self, optics, alpha, device="cuda"):
        import torch
        edges = np.asarray(optics["panel_edges"], float)
        phase = np.asarray(optics["phase_nodes"], float)
        if np.any(np.diff(phase) == 0.0):
            raise ValueError("constant-phase optical panels need the direct-ray evaluator, "
                             "which the GPU path does not provide")
        eff = np.asarray(optics["efficiency_coefficients"], float)
        if eff.shape[1] != 4:
            raise ValueError("the GPU path takes cubic efficiency panels")
        # Lines where the medium table (phase, group, extinction) has a kink:
        # the band ends and the medium nodes inside it.
        medium_nodes = optics.get("medium_nodes_nm")
        if medium_nodes is None:
            raise ValueError("optics must carry the medium nodes (prepare with lighthit_amd.optics)")
        inside = [x for x in medium_nodes if edges[0] < x < edges[-1]]
        major = np.unique(np.asarray([0, len(edges) - 1] + [int(np.searchsorted(edges, x))
                                                              for x in inside], np.int64))
        if not np.allclose(edges[major[1:-1]], inside):
            raise ValueError("medium nodes are not panel edges")
        self.optics = optics
        self.alpha = np.asarray(alpha, float)
        self.major = major.astype(np.int32)
        self._keep = [_dev(edges, torch.float64, device), _dev(phase, torch.float64, device),
                      _dev(optics["group_nodes"], torch.float64, device),
                      _dev(optics["extinction_nodes"], torch.float64, device),
                      _dev(eff, torch.float64, device), _dev(self.major, torch.int32, device),
                      _dev(self.alpha, torch.float64, device)]
        ptrs = [t.data_ptr() for t in self._keep]
        self.struct = _Optics(*ptrs, len(edges) - 1, 4, len(self.major), len(self.alpha),
                              float(phase.min()), float(phase.max()))


class DeviceOMs:
    """Positions, looks, areas and glass radii of a detector (or subset), on the device."""

    def __init__(self, detector, device="cuda"):
        import torch
        radii = glass_radii(detector)
        if radii is None:
            raise ValueError("the first-entry GPU path needs detector.receiver_surface geometry")
        self.count = len(radii)
        area = np.broadcast_to(np.asarray(detector.effective_area_m2, float), (self.count,))
        self._keep = _lib.upload([detector.positions_m, look_directions(detector), area, radii],
                                 device)
        self.struct = _Oms(*[t.data_ptr() for t in self._keep], self.count)


class DeviceElements:
    """Spectral light elements of one source, on the device."""

    @classmethod
    def from_chords(cls, chords):
        """The elements of a :class:`lighthit_amd.compile.DeviceChords` (shares its
        device arrays)."""
        elements = chords.host
        cls._check(elements)
        self = cls.__new__(cls)
        self.count = len(elements)
        self.chords = chords
        self._keep = [chords.start_m, chords.direction, chords.length_m, chords.coefficient0,
                      chords.coefficient2, chords.beta, chords.start_ns, chords.end_ns]
        self.struct = _Elements(*[t.data_ptr() for t in self._keep], self.count)
        return self

    @staticmethod
    def _check(elements):
        if elements.cone_model != "spectral":
            raise ValueError('continuous ballistic spectrum requires spectral LightElements')
        if not np.allclose(elements.coefficient2, -elements.coefficient0 / elements.beta ** 2,
                           rtol=1e-10, atol=1e-20):
            raise ValueError('source must have the physical Cherenkov C0/C2/beta relation')

    def __init__(self, elements, device="cuda"):
        import torch
        self._check(elements)
        self.count = len(elements)
        names = ("start_m", "direction", "length_m", "coefficient0", "coefficient2", "beta",
                 "start_ns", "end_ns")
        # one host-to-device copy for every array
        self._keep = _lib.upload([np.asarray(getattr(elements, n), float) for n in names], device)
        self.struct = _Elements(*[t.data_ptr() for t in self._keep], self.count)


class Order0:
    """Resident order-0 evaluator of one (medium, detector): workspace reused across events."""

    def __init__(self, medium, detector, alpha, *, wavelength_range_nm=None, device="cuda"):
        optics = prepare_optics(medium, detector, wavelength_range_nm)
        self.optics = DeviceOptics(optics, alpha, device)
        self.oms = DeviceOMs(detector, device)
        self.detector = detector
        self.device = device

    def workspace(self, n_elements, stream=None):
        need = _workspace_bytes(self.oms.count, n_elements)
        return _lib.workspace("order0_charge", self.device, need, stream)

    def charge_device(self, elements, out=None, stream=None, oms=None):
        """Enqueue the charge of every OM (or of ``oms``, a :class:`DeviceOMs`);
        returns (charge tensor, workspace)."""
        import torch
        oms = self.oms if oms is None else oms
        if out is None:
            out = torch.empty(oms.count, dtype=torch.float64, device=self.device)
        ws = _lib.workspace("order0_charge", self.device,
                            _workspace_bytes(oms.count, elements.count), stream)
        with _lib.stage("order0_charge"):
            _lib.call("lh_order0_charge", ctypes.byref(self.optics.struct),
                      ctypes.byref(oms.struct), ctypes.byref(elements.struct),
                      ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                      ctypes.c_size_t(ws.numel()), _lib.stream_handle(stream))
        return out, ws

    def charge_noalloc_device(self, elements, out, flags, stream=None):
        """Enqueue the charge through ``lh_order0_charge_noalloc`` into the caller's
        ``out`` (OM,) float64 and ``flags`` (1,) int32 device tensors: no device
        allocation, the library's fixed scratch instead of a workspace."""
        _lib.call("lh_order0_charge_noalloc", ctypes.byref(self.optics.struct),
                  ctypes.byref(self.oms.struct), ctypes.byref(elements.struct),
                  ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(flags.data_ptr()),
                  _lib.stream_handle(stream))
        return out

    def bins_device(self, elements, oms, origins, edges, out=None, stream=None):
        """Enqueue the arrival bins of ``oms`` (a :class:`DeviceOMs`, normally the active
        subset); ``origins`` (OM,) float64 and ``edges`` (bins+1,) float64 device tensors."""
        import torch
        nbins = int(edges.numel()) - 1
        if out is None:
            out = torch.empty((oms.count, nbins), dtype=torch.float64, device=self.device)
        need = _workspace_bytes(oms.count, elements.count, nbins)
        ws = _lib.workspace("order0_bins", self.device, need, stream)
        with _lib.stage("order0_bins"):
            _lib.call("lh_order0_bins", ctypes.byref(self.optics.struct),
                      ctypes.byref(oms.struct), ctypes.byref(elements.struct),
                      ctypes.c_void_p(origins.data_ptr()), ctypes.c_void_p(edges.data_ptr()),
                      nbins, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                      ctypes.c_size_t(ws.numel()), _lib.stream_handle(stream))
        return out, ws

    def bins(self, elements, ids, origins, edges):
        """Bins (len(ids), bins) of OMs ``ids`` (``PreparedBallistic.bins``, numpy)."""
        import torch
        ids = np.asarray(ids, dtype=np.int64)
        origins, edges = np.asarray(origins, float), np.asarray(edges, float)
        if (origins.shape != (len(ids),) or not np.isfinite(origins).all()
                or edges.ndim != 1 or len(edges) < 2 or not np.isfinite(edges).all()
                or np.any(np.diff(edges) <= 0)):
            raise ValueError('invalid time origins or edges')
        device_elements = elements if isinstance(elements, DeviceElements) else DeviceElements(
            elements, self.device)
        oms = DeviceOMs(self.detector.subset(ids), self.device)
        out, ws = self.bins_device(device_elements, oms,
                                   *_lib.upload([origins, edges], self.device))
        torch.cuda.synchronize()
        check_flags(ws)
        return out.cpu().numpy()

    #: Largest batch (events, element chunks) replayed from a graph; larger ones
    #: launch the same chain directly.
    GRAPH_EVENTS, GRAPH_CHUNKS = 16, 4096

    def oms_of(self, ids):
        """The :class:`DeviceOMs` of OMs ``ids`` of the detector (all OMs: None)."""
        return self.oms if ids is None else DeviceOMs(self.detector.subset(np.asarray(ids)),
                                                      self.device)

    def charge_batch_device(self, events, *, graph=False, stream=None, oms=None):
        """Enqueue the charges of every OM for several events in one fused launch
        (``lh_order0_charge_batch``: one persistent kernel over the (event, OM,
        chunk) items of all of them); returns ((events, OM) charge tensor,
        workspace). Each event's row equals its :meth:`charge_device` result bit
        for bit (the same items, sums in the same order).

        With ``graph`` the launch chain is captured once per capacity (events
        rounded up to a power of two, element chunks to 256, 1024 or 4096) as a
        HIP graph and replayed; only the small table of this batch's events is
        written before the replay. ``oms`` (a :class:`DeviceOMs`, default all)
        evaluates those OMs only, without a graph."""
        import torch
        events = list(events)
        n = len(events)
        graph = graph and oms is None
        oms = self.oms if oms is None else oms
        if n == 0:
            return torch.empty((0, oms.count), dtype=torch.float64, device=self.device), None
        counts = (ctypes.c_int * n)(*[int(e.count) for e in events])
        chunks = int(_query("lh_order0_batch_chunks", counts, n))
        graph = graph and n <= self.GRAPH_EVENTS and chunks <= self.GRAPH_CHUNKS
        if graph:
            # capacities in coarse steps, so that few graphs serve every batch;
            # all of them share one workspace of the largest capacity (a
            # process replays them one at a time, in stream order)
            max_events = _bucket(n, 2)
            max_chunks = 256
            while max_chunks < chunks:
                max_chunks *= 4
            if getattr(self, "_graph_ws", None) is None:
                self._graph_ws = torch.empty(int(_query(
                    "lh_order0_batch_workspace_bytes", oms.count, self.GRAPH_EVENTS,
                    self.GRAPH_CHUNKS)), dtype=torch.uint8, device=self.device)
                self._graphs = {}
            ws = self._graph_ws
            captured = self._graphs.get((max_events, max_chunks))
        else:
            max_events, max_chunks = n, chunks
            need = int(_query("lh_order0_batch_workspace_bytes", oms.count, max_events,
                              max_chunks))
            ws = _lib.workspace("order0_batch", self.device, need, stream)
        out = torch.empty((n, oms.count), dtype=torch.float64, device=self.device)
        table = (_Elements * n)(*[e.struct for e in events])
        outs = (ctypes.c_void_p * n)(*[out[k].data_ptr() for k in range(n)])
        handle = _lib.stream_handle(stream)
        with _lib.stage("order0_charge"):
            _lib.call("lh_order0_batch_set", ctypes.byref(oms.struct), table, outs, n,
                      max_events, max_chunks, ctypes.c_void_p(ws.data_ptr()),
                      ctypes.c_size_t(ws.numel()), handle)
            args = (ctypes.byref(self.optics.struct), ctypes.byref(oms.struct), max_events,
                    max_chunks, ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(ws.numel()))
            if not graph:
                _lib.call("lh_order0_charge_batch", *args, handle)
            else:
                if captured is None:
                    captured = self._capture(args, ws, stream)
                captured[0].replay(stream)
        return out, ws

    def _capture(self, args, ws, stream):
        """Capture the fused charge chain of one capacity as a HIP graph."""
        if not hasattr(self, "_graphs"):
            self._graphs = {}
        graph = _lib.Graph.capture(
            lambda side: _lib.call("lh_order0_charge_batch", *args, _lib.stream_handle(side)),
            self.device)
        self._graphs[(args[2], args[3])] = (graph, ws)
        return self._graphs[(args[2], args[3])]

    def charge_batch(self, events, *, graph=False):
        """Charges per OM (numpy, one row per event) of several events in one fused
        launch, raising the reference's errors (the event at fault is found by
        evaluating the events one by one)."""
        import torch
        device_events = [e if isinstance(e, DeviceElements) else DeviceElements(e, self.device)
                         for e in events]
        out, ws = self.charge_batch_device(device_events, graph=graph)
        torch.cuda.synchronize()
        if ws is not None:
            try:
                check_flags(ws)
            except ValueError:
                return np.stack([self.charge(e) for e in device_events])
        return out.cpu().numpy()

    def charge(self, elements, ids=None):
        """Charge per OM (numpy), raising the reference's errors; with ``ids``
        the charges of those OMs only (the others are not evaluated)."""
        import torch
        device_elements = elements if isinstance(elements, DeviceElements) else DeviceElements(
            elements, self.device)
        out, ws = self.charge_device(device_elements, oms=self.oms_of(ids))
        torch.cuda.synchronize()
        check_flags(ws)
        return out.cpu().numpy()


def _workspace_bytes(n_oms, n_elements, n_bins=None):
    lib = _lib.load()
    name = "lh_order0_workspace_bytes" if n_bins is None else "lh_order0_bins_workspace_bytes"
    fn = getattr(lib, name)
    fn.restype, fn.argtypes = _lib.SIGNATURES[name]
    args = (int(n_oms), int(n_elements)) + (() if n_bins is None else (int(n_bins),))
    return int(fn(*args))


def check_flags(workspace):
    """Raise the reference's error for flags an order-0 launch left in its workspace."""
    import torch
    flags = int(workspace[:4].view(dtype=torch.int32)[0].item())
    if flags & FLAG_GLASS:
        raise ValueError('emitting segment intersects the glass boundary')
    if flags & FLAG_UNSUPPORTED:
        raise ValueError("an element line passes within the glass radius of an OM centre; "
                         "the direct-ray evaluator for such pairs is not on the GPU path")
