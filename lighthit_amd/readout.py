"""Stage (c): time bins of order >=2 light from its spectrum, on the GPU.

Host mirror of the reference's readout (``readout.py:9-72``) and of the
readout step of the shower route (``shower/api.py:446-478``). The per-grid
work — bin integrals ``K`` on the readout grid and the not-a-knot spline
``S`` from the table grid onto it — is folded once into ``M = K S`` (FP64 on
the host, stored FP32 on the device); each event is then one MFMA launch of
``lh_readout_bins`` with the demodulation by the OM time origins fused in.
"""
from __future__ import annotations

import warnings

import numpy as np

from . import _lib


def frequency_weights(omega):
    """Trapezoid weights of a grid starting at zero (``readout.py:9-15``)."""
    omega = np.asarray(omega, float)
    if (omega.ndim != 1 or len(omega) < 2 or not np.isfinite(omega).all()
            or omega[0] != 0 or np.any(np.diff(omega) <= 0)):
        raise ValueError("Time inversion requires at least two increasing frequencies starting at zero")
    step = np.diff(omega)
    return np.r_[step[0] / 2, (step[:-1] + step[1:]) / 2, step[-1] / 2]


def _check_edges(edges):
    edges = np.asarray(edges, float)
    if edges.ndim != 1 or len(edges) < 2 or not np.isfinite(edges).all() or np.any(np.diff(edges) <= 0):
        raise ValueError("Time edges must be finite and strictly increasing")
    return edges


def bin_kernel(omega, edges_ns, sigma_ns=0.0):
    """Complex (bin, frequency) matrix of ``inverse_bins`` (``readout.py:34-39``).

    Row ``b`` integrates ``exp(-i w t)`` exactly over bin ``b`` (a sinc of the
    bin width times the phase of its centre), weighted by the trapezoid weights
    and an optional Gaussian time smearing, divided by pi.
    """
    omega = np.asarray(omega, float)
    edges = _check_edges(edges_ns)
    if not np.isfinite(sigma_ns) or sigma_ns < 0:
        raise ValueError("sigma_ns must be finite and nonnegative")
    w = frequency_weights(omega)
    width = np.diff(edges)[:, None]
    centre = ((edges[:-1] + edges[1:]) / 2)[:, None]
    return (width * np.sinc(width * omega / (2 * np.pi)) * np.exp(-1j * centre * omega)
            * (w * np.exp(-0.5 * (omega * sigma_ns) ** 2) / np.pi))


def spline_matrix(x, at):
    """Matrix ``S`` with ``S @ y`` = not-a-knot cubic spline of ``(x, y)`` at ``at``.

    The spline is written per interval in Hermite form with unknown knot slopes
    ``s``. Interior knots impose a continuous second derivative; the two knots
    next to the ends impose a continuous third derivative (not-a-knot), i.e.
    equal cubic coefficients on their two intervals. Values outside ``x`` use
    the end cubics. Needs at least four knots.
    """
    x = np.asarray(x, float)
    at = np.asarray(at, float)
    n = len(x)
    if n < 4 or np.any(np.diff(x) <= 0):
        raise ValueError("the spline needs at least four increasing knots")
    h = np.diff(x)
    eye = np.eye(n)
    slope = (eye[1:] - eye[:-1]) / h[:, None]          # m_i as rows acting on y
    A = np.zeros((n, n))
    B = np.zeros((n, n))
    # cubic coefficient of interval i: (s_i + s_{i+1} - 2 m_i) / h_i^2
    for i in (0, n - 3):
        # equal cubic coefficients on intervals i and i+1
        A[0 if i == 0 else n - 1, i] += 1 / h[i] ** 2
        A[0 if i == 0 else n - 1, i + 1] += 1 / h[i] ** 2 - 1 / h[i + 1] ** 2
        A[0 if i == 0 else n - 1, i + 2] -= 1 / h[i + 1] ** 2
        B[0 if i == 0 else n - 1] = 2 * slope[i] / h[i] ** 2 - 2 * slope[i + 1] / h[i + 1] ** 2
    for i in range(1, n - 1):
        # second derivative continuous at x_i:
        # left end of interval i:   (6 m_i - 4 s_i - 2 s_{i+1}) / h_i
        # right end of interval i-1: (-6 m_{i-1} + 2 s_{i-1} + 4 s_i) / h_{i-1}
        A[i, i - 1] = 2 / h[i - 1]
        A[i, i] = 4 / h[i - 1] + 4 / h[i]
        A[i, i + 1] = 2 / h[i]
        B[i] = 6 * slope[i - 1] / h[i - 1] + 6 * slope[i] / h[i]
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):                   # the same rounding in every process
        slopes = np.linalg.solve(A, B)                  # s = slopes @ y
    index = np.clip(np.searchsorted(x, at, side="right") - 1, 0, n - 2)
    t = (at - x[index])[:, None]
    hi = h[index][:, None]
    m = slope[index]
    s0, s1 = slopes[index], slopes[index + 1]
    c2 = (3 * m - 2 * s0 - s1) / hi
    c3 = (s0 + s1 - 2 * m) / hi ** 2
    return eye[index] + s0 * t + c2 * t ** 2 + c3 * t ** 3


def fixed_order_product(a, b):
    """``a @ b`` summed over the inner index in order, one rank-1 update at a
    time: the same bits whatever the BLAS library and its thread count (a
    threaded matrix product rounds differently from one process to the next)."""
    out = np.zeros((a.shape[0], b.shape[1]), np.result_type(a, b))
    for k in range(a.shape[1]):
        out += a[:, k, None] * b[None, k, :]
    return out


class ReadoutBasis:
    """Folded readout basis of one (table grid, readout grid, bin edges) triple.

    ``matrix`` is the FP64 complex ``M = K S`` (bins x table frequencies);
    ``device`` the FP32 ``[Re M | -Im M]`` handed to ``lh_readout_bins``.
    """

    def __init__(self, omega, edges_ns, readout_omega=None, sigma_ns=0.0, device="cuda"):
        import torch
        self.omega = np.asarray(omega, float)
        self.edges = _check_edges(edges_ns)
        frequency_weights(self.omega)
        grid = self.omega
        spline = np.eye(len(self.omega))
        if readout_omega is not None and len(self.omega) > 1:
            readout = np.asarray(readout_omega, float)
            same = len(readout) == len(self.omega) and np.allclose(readout, self.omega)
            if not same and np.isclose(readout[-1], self.omega[-1], rtol=0, atol=1e-12):
                spline = spline_matrix(self.omega, readout)
                grid = readout
        self.grid = grid
        self.matrix = fixed_order_product(bin_kernel(grid, self.edges, sigma_ns), spline)
        packed = np.concatenate([self.matrix.real, -self.matrix.imag], axis=1)
        self.device = torch.as_tensor(np.ascontiguousarray(packed, np.float32), device=device)
        self.omega_device = torch.as_tensor(self.omega, dtype=torch.float64, device=device)

    @property
    def bins(self):
        return len(self.edges) - 1

    def period_ns(self):
        return 2 * np.pi / float(np.max(np.diff(self.grid))) if len(self.grid) > 1 else np.inf

    def split(self):
        """Device FP16 ``(hi, lo, row_scale)`` of ``[Re M | -Im M]`` for the split-precision
        readout: every row scaled by a power of two so its largest value lies in
        [2^13, 2^14), hi its FP16 rounding and lo the FP16 rounding of the rest."""
        if getattr(self, "_split", None) is None:
            import torch
            packed = np.concatenate([self.matrix.real, -self.matrix.imag], axis=1)
            peak = np.abs(packed).max(axis=1)
            exponent = np.where(peak > 0, np.floor(np.log2(np.where(peak > 0, peak, 1.0))), 0.0)
            row_scale = np.ldexp(1.0, (13 - exponent).astype(int))
            scaled = packed * row_scale[:, None]
            hi = scaled.astype(np.float16)
            lo = (scaled - hi.astype(np.float64)).astype(np.float16)
            device = self.device.device
            self._split = (torch.as_tensor(hi, device=device), torch.as_tensor(lo, device=device),
                           torch.as_tensor(row_scale.astype(np.float32), device=device))
        return self._split

    def apply(self, spectrum, origins=None, out=None, stream=None, precision="fp32"):
        """Bins (observation, bin) float32 of device spectra (frequency, observation) complex64.

        ``origins`` (observation,) float64 demodulates absolute-time spectra.
        ``precision`` "fp32" (FP32 MFMA) or "split" (FP16 hi + lo MFMA).
        """
        import torch
        if spectrum.dtype != torch.complex64 or spectrum.dim() != 2:
            raise ValueError("spectrum must be a (frequency, observation) complex64 tensor")
        nfreq, nobs = spectrum.shape
        if nfreq != len(self.omega):
            raise ValueError("spectrum and readout basis frequency grids differ")
        if origins is not None and (origins.dtype != torch.float64 or origins.shape != (nobs,)):
            raise ValueError("origins must be a float64 tensor with one value per observation")
        if out is None:
            out = torch.empty((nobs, self.bins), dtype=torch.float32, device=spectrum.device)
        if precision not in ("fp32", "split"):
            raise ValueError("precision must be 'fp32' or 'split'")
        with _lib.stage("readout"):
            if precision == "split":
                hi, lo, row_scale = self.split()
                _lib.call("lh_readout_bins_split", _lib.ptr(hi), _lib.ptr(lo),
                          _lib.ptr(row_scale), _lib.ptr(self.omega_device), _lib.ptr(spectrum),
                          _lib.ptr(origins), self.bins, nfreq, nobs, _lib.ptr(out),
                          _lib.stream_handle(stream))
            else:
                _lib.call("lh_readout_bins", _lib.ptr(self.device), _lib.ptr(self.omega_device),
                          _lib.ptr(spectrum), _lib.ptr(origins), self.bins, nfreq, nobs,
                          _lib.ptr(out), _lib.stream_handle(stream))
        return out


def _run(omega, spectrum, edges_ns, readout_omega, sigma_ns):
    import torch
    spectrum = np.asarray(spectrum, complex)
    omega = np.asarray(omega, float)
    if spectrum.ndim != 2 or spectrum.shape[0] != len(omega) or not np.isfinite(spectrum).all():
        raise ValueError("spectrum must have shape (frequency, observation)")
    basis = ReadoutBasis(omega, edges_ns, readout_omega, sigma_ns)
    device = torch.as_tensor(np.ascontiguousarray(spectrum, np.complex64), device="cuda")
    out = basis.apply(device)
    return out.double().cpu().numpy(), basis


def inverse_bins(omega, spectrum, edges_ns, sigma_ns=0.0):
    """Drop-in for ``readout.inverse_bins`` (``readout.py:18-40``); (observation, bin)."""
    _check_edges(edges_ns)
    if not np.isfinite(sigma_ns) or sigma_ns < 0:
        raise ValueError("sigma_ns must be finite and nonnegative")
    return _run(omega, spectrum, edges_ns, None, sigma_ns)[0]


def readout_bins(omega, spectrum, edges_ns, readout_omega=None):
    """Drop-in for ``readout.readout_bins`` (``readout.py:43-72``)."""
    edges = _check_edges(edges_ns)
    bins, basis = _run(omega, spectrum, edges, readout_omega, 0.0)
    period = basis.period_ns()
    if edges[-1] - edges[0] > period * (1 + 1e-9):
        warnings.warn(f"time window of {edges[-1] - edges[0]:g} ns is longer than the "
                      f"{period:g} ns period of the frequency grid; the bins hold repeated "
                      "copies of the waveform (set KernelConfig.readout_omega_per_ns)",
                      RuntimeWarning, stacklevel=2)
    return bins
