"""ctypes binding of ``liblighthit_amd.so`` (the gfx950 C-ABI).

The library is loaded from the package directory, never from a cache, and a
missing library is an error: there is no CPU fallback for any entry point.
Device buffers are torch tensors on a ROCm device; only their addresses cross
the boundary.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIBRARY = Path(__file__).resolve().parent / "liblighthit_amd.so"

LH_OK, LH_ERR_ARGUMENT, LH_ERR_HIP = 0, 1, 2

_p = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_f = ctypes.c_float
_i64 = ctypes.c_int64

#: Entry points and their argument types, as declared in include/lighthit_amd.h.
SIGNATURES = {
    "lh_last_error": (ctypes.c_char_p, []),
    "lh_clear_error": (None, []),
    "lh_abi_version": (_i, []),
    "lh_readout_bins": (_i, [_p, _p, _p, _p, _i, _i, _i, _p, _p]),
    "lh_readout_bins_split": (_i, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _p, _p]),
    "lh_timing_enable": (_i, [_i]),
    "lh_timing_read": (_i, [ctypes.c_char_p, ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_longlong)]),
    "lh_timing_reset": (_i, []),
    "lh_graph_begin": (_i, [_p]),
    "lh_graph_end": (_i, [_p, ctypes.POINTER(ctypes.c_void_p)]),
    "lh_graph_launch": (_i, [_p, _p]),
    "lh_graph_destroy": (_i, [_p]),
}

_lib = None


def load():
    """The loaded library (raises ``RuntimeError`` when it was not built)."""
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise RuntimeError(
                f"{LIBRARY} is missing: build it with `python -m lighthit_amd.build` "
                "(lighthit_amd has no CPU fallback)")
        lib = ctypes.CDLL(str(LIBRARY))
        for name in ("lh_last_error", "lh_clear_error", "lh_abi_version"):
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = SIGNATURES[name]
        _lib = lib
    return _lib


_bound = set()


def call(name, *args):
    """Call entry point ``name``; map a failed status to the reference's exceptions."""
    lib = load()
    fn = getattr(lib, name)
    if name not in _bound:
        restype, argtypes = SIGNATURES[name]
        fn.restype, fn.argtypes = restype, argtypes
        _bound.add(name)
    status = fn(*args)
    if status != LH_OK:
        message = lib.lh_last_error().decode(errors="replace")
        lib.lh_clear_error()
        if status == LH_ERR_ARGUMENT:
            raise ValueError(message)
        raise RuntimeError(message)


def ptr(tensor):
    """Device address of a torch tensor (None -> NULL)."""
    if tensor is None:
        return None
    if not tensor.is_cuda:
        raise ValueError("lighthit_amd entry points take device tensors")
    if not tensor.is_contiguous():
        raise ValueError("lighthit_amd entry points take contiguous tensors")
    return ctypes.c_void_p(tensor.data_ptr())


def stream_handle(stream=None):
    """hipStream_t of ``stream`` (default: torch's current stream)."""
    import torch
    stream = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(stream.cuda_stream)


def upload(arrays, device, dtype=None):
    """Device copies of several host arrays of one dtype (default float64) in
    one host-to-device copy: views into one buffer, each piece starting on a
    256-byte boundary. Small per-event uploads then cost one copy, not one each."""
    import numpy as np
    import torch
    dtype = torch.float64 if dtype is None else dtype
    np_dtype = torch.empty((), dtype=dtype).numpy().dtype
    arrays = [np.ascontiguousarray(a, dtype=np_dtype) for a in arrays]
    align = max(1, 256 // np_dtype.itemsize)
    offsets, at = [], 0
    for a in arrays:
        offsets.append(at)
        at += -(-a.size // align) * align
    host = np.zeros(max(at, 1), dtype=np_dtype)
    for a, off in zip(arrays, offsets):
        host[off:off + a.size] = a.ravel()
    buf = torch.from_numpy(host).to(device)
    return [buf.narrow(0, off, a.size).view(a.shape) for a, off in zip(arrays, offsets)]


_workspaces = {}
_workspaces_lock = threading.Lock()


def workspace(tag, device, nbytes, stream=None):
    """A reusable device buffer of at least ``nbytes`` for ``tag`` on ``stream``
    (default: the current stream). Work on one stream runs in order, so a
    buffer per (tag, device, stream) is never in use by two calls at once;
    callers on different streams get different buffers."""
    import torch
    device = torch.device(device)
    if stream is None:
        stream = torch.cuda.current_stream(device)
    key = (tag, str(device), int(stream.cuda_stream))
    with _workspaces_lock:
        buf = _workspaces.get(key)
        if buf is None or buf.numel() < nbytes:
            # the smaller buffer goes back to the allocator first (its reuse is
            # ordered after the work already queued on this stream)
            _workspaces.pop(key, None)
            del buf
            buf = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
            _workspaces[key] = buf
    return buf


def release_workspaces(device=None):
    """Drop the cached workspaces (of ``device``, or all); their memory returns
    to the allocator."""
    import torch
    with _workspaces_lock:
        for key in [k for k in _workspaces
                    if device is None or k[1] == str(torch.device(device))]:
            del _workspaces[key]


class StageClock:
    """HIP events around the library's stages on the launch stream.

    While a clock is active (``with StageClock() as clock``), every wrapped
    stage records a start and an end event on the current stream;
    :meth:`milliseconds` sums the intervals per stage after synchronising.
    """

    active = None

    def __init__(self):
        self.marks = []

    def __enter__(self):
        StageClock.active = self
        return self

    def __exit__(self, *exc):
        StageClock.active = None
        return False

    def milliseconds(self):
        import torch
        torch.cuda.synchronize()
        out = {}
        for name, start, end in self.marks:
            out[name] = out.get(name, 0.0) + start.elapsed_time(end)
        return out

    def counts(self):
        out = {}
        for name, _, _ in self.marks:
            out[name] = out.get(name, 0) + 1
        return out


class stage:
    """Context manager recording one stage interval on the active clock (if any)."""

    def __init__(self, name):
        self.name = name
        self.start = None

    def __enter__(self):
        if StageClock.active is not None:
            import torch
            self.start = torch.cuda.Event(enable_timing=True)
            self.start.record()
        return self

    def __exit__(self, *exc):
        clock = StageClock.active
        if clock is not None and self.start is not None:
            import torch
            end = torch.cuda.Event(enable_timing=True)
            end.record()
            clock.marks.append((self.name, self.start, end))
        return False


class KernelTimer:
    """HIP events that the library records around its dominant kernels.

    ``with KernelTimer() as timer:`` turns the library's kernel timer on (and
    clears earlier marks); ``timer.read(name)`` returns ``(total_ms, launches)``
    of every launch of kernel ``name`` inside the block.
    """

    def __enter__(self):
        call("lh_timing_reset")
        call("lh_timing_enable", 1)
        return self

    def __exit__(self, *exc):
        call("lh_timing_enable", 0)
        return False

    @staticmethod
    def read(name):
        total, launches = ctypes.c_double(), ctypes.c_longlong()
        call("lh_timing_read", name.encode(), ctypes.byref(total), ctypes.byref(launches))
        return total.value, launches.value

    @staticmethod
    def reset():
        call("lh_timing_reset")


class Graph:
    """A captured launch sequence of this library (``lh_graph_*``), replayed
    with :meth:`replay` on torch's current stream of its device.

    ``Graph.capture(enqueue, device)`` runs ``enqueue(stream)`` on a side
    stream under capture (the enqueued launches are recorded, not run) and
    instantiates the result. Unlike ``torch.cuda.graph`` it neither
    synchronises the device nor empties torch's allocator cache, so a serving
    process can capture while other work is in flight; ``enqueue`` must
    allocate nothing (its buffers exist before the capture)."""

    def __init__(self, handle, device):
        self._handle = handle
        self.device = device

    @classmethod
    def capture(cls, enqueue, device):
        import torch
        side = torch.cuda.Stream(device=device)
        side.wait_stream(torch.cuda.current_stream(device))
        handle = ctypes.c_void_p()
        raw = ctypes.c_void_p(side.cuda_stream)
        call("lh_graph_begin", raw)
        try:
            enqueue(side)
        finally:
            call("lh_graph_end", raw, ctypes.byref(handle))
        return cls(handle, device)

    def replay(self, stream=None):
        call("lh_graph_launch", self._handle, stream_handle(stream))

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value and _lib is not None:
            try:
                call("lh_graph_destroy", handle)
            except Exception:
                pass
