"""Order >=2 kernel pack and its band fold (table preparation).

Mirrors ``BandFold``, ``fold_bundle``, ``zero_frequency``, ``save_folded`` and
``load_folded`` (``band_fold.py:62-240``) and the pack store
``Ge2KernelStore.save``/``load``/``bundle`` (``shower_basis.py:191-349``,
schema ``lighthit/ge2-kernel-pack/1``), so packs and folds written by the
reference load unchanged. The node-band choice is a small least-squares
problem on a beta grid (host FP64, once per medium and detector); the re-fit
of the folded radial kernels runs on the GPU (``lh_fold_bundle``).
"""
from __future__ import annotations

import ctypes
import json
from dataclasses import dataclass
from itertools import combinations
from math import comb
from pathlib import Path

import numpy as np

from . import _lib

FORMAT = "lighthit/ge2-band-fold/1"
PACK_SCHEMA = "lighthit/ge2-kernel-pack/1"
#: Largest number of band subsets searched exhaustively (``band_fold.py:47``).
EXHAUSTIVE_SUBSETS = 10_000


class _FoldInput(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("coefficients", "log_r", "new_log_r", "omega",
                                               "absorption", "speed", "mixing", "inverse")] + [
        ("absorption_ref", ctypes.c_double), ("speed_ref", ctypes.c_double)] + [
        (n, ctypes.c_int) for n in ("bands", "intervals", "degree", "pairs", "freq", "rank",
                                    "subdivide")]


_lib.SIGNATURES["lh_fold_bundle"] = (ctypes.c_int, [ctypes.POINTER(_FoldInput), ctypes.c_void_p,
                                                    ctypes.c_void_p])


def _legendre_table(degree, x):
    """``P_l(x)`` for l = 0..degree (``band_fold.py:50-58``)."""
    out = np.empty((degree + 1,) + np.shape(x))
    out[0] = 1.0
    if degree:
        out[1] = x
    for ell in range(2, degree + 1):
        out[ell] = ((2 * ell - 1) * x * out[ell - 1] - (ell - 1) * out[ell - 2]) / ell
    return out


@dataclass(frozen=True)
class BandFold:
    """Node bands and the per-degree interpolation ``L[l, lambda, j]`` (``band_fold.py:62-152``)."""
    bands: tuple
    interpolation: np.ndarray
    scale: np.ndarray
    phase_index: np.ndarray
    residual_beta: np.ndarray
    residual: np.ndarray
    beta_design_min: float

    @classmethod
    def of(cls, kernel, *, rank=3, degree=None, beta_design_min=0.8, bands=None):
        """Choose ``rank`` node bands minimising the worst relative column residual of
        ``F_l(lambda; beta) = a_lambda Q P_l(1/(beta n))`` for beta above
        ``beta_design_min`` (exhaustive up to ``EXHAUSTIVE_SUBSETS`` subsets, else
        greedy), unless ``bands`` is given."""
        degree = int(kernel.config.source_degree if degree is None else degree)
        wavelength = np.asarray(kernel.wavelength.wavelength_nm, float)
        weight = np.asarray(kernel.wavelength.weight_nm, float)
        efficiency = np.asarray(kernel.detector.spectral_weight(wavelength), float)
        phase = np.asarray(kernel.medium.sample(wavelength)["phase_index"], float)
        scale = weight * efficiency
        nband = len(wavelength)
        if not 1 <= rank <= nband:
            raise ValueError("rank must lie between 1 and the number of bands")
        epsilon = np.r_[0.0, np.geomspace(1e-7, phase.max() - 1.0, 800)]
        beta = 1.0 / (1.0 + epsilon)
        cosine = 1.0 / (beta[None, :] * phase[:, None])
        yield_ = (scale / wavelength ** 2)[:, None] * np.clip(1.0 - cosine ** 2, 0.0, None)
        columns = _legendre_table(degree, np.clip(cosine, -1.0, 1.0)) * yield_[None]
        norm = np.maximum(np.linalg.norm(yield_, axis=0), 1e-300)
        design = beta > beta_design_min

        def residuals(subset):
            nodes = columns[:, list(subset), :]
            mixing = columns @ np.linalg.pinv(nodes)
            remainder = columns - mixing @ nodes
            return mixing, np.linalg.norm(remainder, axis=1) / norm[None, :]

        def worst(subset):
            return float(residuals(subset)[1][:, design].max())

        if bands is None and comb(nband, rank) <= EXHAUSTIVE_SUBSETS:
            bands = min(combinations(range(nband), rank), key=worst)
        elif bands is None:
            bands = ()
            for _ in range(rank):
                bands = min((tuple(sorted(bands + (band,))) for band in range(nband)
                             if band not in bands), key=worst)
        bands = tuple(int(b) for b in bands)
        mixing, residual = residuals(bands)
        return cls(bands, np.ascontiguousarray(mixing), scale, phase, beta,
                   residual.max(axis=0), float(beta_design_min))

    def source_residual(self, beta, weight):
        """Light-weighted relative band-interpolation residual of a source."""
        beta = np.asarray(beta, float)
        weight = np.asarray(weight, float)
        order = np.argsort(self.residual_beta)
        values = np.interp(beta, self.residual_beta[order], self.residual[order])
        return float(np.sum(weight * values) / max(np.sum(weight), 1e-300))

    def metadata(self):
        design = self.residual_beta > self.beta_design_min
        return dict(bands=list(self.bands), rank=len(self.bands),
                    beta_design_min=self.beta_design_min,
                    worst_residual_design=float(self.residual[design].max()),
                    worst_residual_all=float(self.residual.max()))


def refit_geometry(log_r, subdivide):
    """Subdivided radial grid and the inverse Vandermonde matrix of each new interval
    (node values at 0, 1/3, 2/3, 1 of the interval -> cubic coefficients)."""
    log_r = np.asarray(log_r, float)
    nint = len(log_r) - 1
    parts = np.linspace(0.0, 1.0, int(subdivide) + 1)
    new_log_r = np.concatenate([log_r[j] + parts[:-1] * (log_r[j + 1] - log_r[j])
                                for j in range(nint)] + [log_r[-1:]])
    nodes = np.array([0.0, 1 / 3, 2 / 3, 1.0])
    inverse = np.stack([np.linalg.inv(np.vander(nodes * (new_log_r[k + 1] - new_log_r[k]), 4))
                        for k in range(len(new_log_r) - 1)])
    return new_log_r, inverse


def fold_bundle(bundle, fold, *, subdivide=1, device="cuda"):
    """Folded bundle with ``len(fold.bands)`` virtual bands (``band_fold.py:155-208``);
    the coefficients are a complex64 device tensor."""
    import torch
    coefficients = bundle["coefficients"]
    nband, nint, ndeg, npair, _, nfreq = coefficients.shape
    if fold.interpolation.shape[:2] != (ndeg, nband):
        raise ValueError("band fold and kernel bundle differ in degree or bands")
    if bundle.get("phase") != "flight":
        raise ValueError("band folding requires flight-phase radial splines")
    log_r = np.asarray(bundle["log_r"], float)
    omega = np.asarray(bundle["omega"], float)
    absorption = np.asarray(bundle["absorption"], float)
    speed = np.asarray(bundle["speed"], float)
    absorption_ref, speed_ref = float(absorption.min()), float(np.mean(speed))
    rank = len(fold.bands)
    new_log_r, inverse = refit_geometry(log_r, subdivide)

    def up(a, dtype=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)
    if isinstance(coefficients, torch.Tensor):
        source = coefficients.to(device=device, dtype=torch.complex64).contiguous()
    else:
        source = up(np.asarray(coefficients, np.complex64), torch.complex64)
    keep = [up(log_r), up(new_log_r), up(omega), up(absorption), up(speed),
            up(fold.interpolation), up(inverse)]
    out = torch.empty((rank, nint * int(subdivide), ndeg, npair, 4, nfreq), dtype=torch.complex64,
                      device=device)
    descriptor = _FoldInput(source.data_ptr(), *[t.data_ptr() for t in keep], absorption_ref,
                            speed_ref, nband, nint, ndeg - 1, npair, nfreq, rank, int(subdivide))
    _lib.call("lh_fold_bundle", ctypes.byref(descriptor), ctypes.c_void_p(out.data_ptr()),
              _lib.stream_handle())
    result = {key: bundle[key] for key in ("degree", "acceptance_degree", "pair_lambda",
                                          "pair_mu", "omega", "phase")}
    result.update(coefficients=out, log_r=new_log_r, absorption=np.full(rank, absorption_ref),
                  speed=np.full(rank, speed_ref), band_fold=fold.metadata(),
                  subdivide=int(subdivide))
    return result


def zero_frequency(folded):
    """The omega = 0 part of a folded bundle (``band_fold.py:211-216``)."""
    if float(folded["omega"][0]) != 0.0:
        raise ValueError("the folded bundle does not start at omega = 0")
    return dict(folded, coefficients=folded["coefficients"][..., :1].contiguous()
                if hasattr(folded["coefficients"], "contiguous")
                else np.ascontiguousarray(folded["coefficients"][..., :1]),
                omega=np.asarray(folded["omega"][:1]))


def save_folded(folded, directory):
    """``folded.npy`` + ``folded.json`` (``band_fold.py:219-226``)."""
    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    coefficients = folded["coefficients"]
    if hasattr(coefficients, "cpu"):
        coefficients = coefficients.cpu().numpy()
    np.save(directory / "folded.npy", np.asarray(coefficients))
    meta = {key: (np.asarray(value).tolist() if isinstance(value, np.ndarray) else value)
            for key, value in folded.items() if key != "coefficients"}
    (directory / "folded.json").write_text(json.dumps(dict(format=FORMAT, **meta)))


def load_folded(directory, *, mmap=True):
    """Read a folded bundle (``band_fold.py:229-240``)."""
    directory = Path(directory)
    meta = json.loads((directory / "folded.json").read_text())
    if meta.pop("format", None) != FORMAT:
        raise ValueError(f"{directory} is not a {FORMAT} pack")
    for key in ("pair_lambda", "pair_mu"):
        meta[key] = np.asarray(meta[key], np.int64)
    for key in ("omega", "log_r", "absorption", "speed"):
        meta[key] = np.asarray(meta[key], float)
    coefficients = np.load(directory / "folded.npy", mmap_mode="r" if mmap else None)
    return dict(meta, coefficients=coefficients)


class KernelPack:
    """Per-wavelength order >=2 radial splines (``Ge2KernelStore``, ``shower_basis.py:191``).

    Bundles are keyed by (wavelengths, degree, frequency indices) and hold
    ``coefficients (bands, intervals, degree+1, pairs, 4, freq)``, ``log_r``,
    ``omega``, ``absorption``, ``speed``, ``pair_lambda``, ``pair_mu``,
    ``degree``, ``acceptance_degree`` and ``phase``.
    """

    def __init__(self, dtype="complex64"):
        self.dtype = np.dtype(dtype)
        self._bundles = {}

    def add(self, wavelengths, indices, bundle):
        key = (tuple(float(w) for w in wavelengths), int(bundle["degree"]),
               tuple(int(i) for i in indices))
        self._bundles[key] = bundle
        return key

    def bundle(self, wavelengths, degree, indices):
        """The bundle of these wavelengths and degree at frequency ``indices`` (a stored
        bundle or a frequency subset of one, ``shower_basis.py:212-243``)."""
        wavelengths = tuple(float(w) for w in wavelengths)
        requested = np.asarray(indices, np.int64)
        for (waves, deg, stored), prepared in self._bundles.items():
            if deg != int(degree) or not np.allclose(waves, wavelengths, rtol=0, atol=1e-9):
                continue
            stored = np.asarray(stored, np.int64)
            locations = np.searchsorted(stored, requested)
            if np.any(locations >= len(stored)) or not np.array_equal(stored[locations],
                                                                       requested):
                continue
            return dict(prepared, coefficients=np.ascontiguousarray(
                np.asarray(prepared["coefficients"])[..., locations]),
                        omega=np.asarray(prepared["omega"])[locations])
        raise ValueError("the order >=2 pack has no bundle for these wavelengths, degree "
                         "and frequencies")

    def save(self, directory):
        """``<n>.npy`` + ``pack.json`` (``shower_basis.py:317-330``)."""
        directory = Path(directory)
        directory.mkdir(parents=True, exist_ok=True)
        index = []
        for number, (key, bundle) in enumerate(self._bundles.items()):
            np.save(directory / f"{number}.npy", np.asarray(bundle["coefficients"]))
            meta = {name: (value.tolist() if isinstance(value, np.ndarray) else value)
                    for name, value in bundle.items()
                    if name != "coefficients" and not name.startswith("_")}
            index.append({"file": f"{number}.npy", "key": [list(key[0]), key[1], list(key[2])],
                          "meta": meta})
        (directory / "pack.json").write_text(json.dumps(
            {"schema": PACK_SCHEMA, "dtype": self.dtype.name, "bundles": index}))

    @classmethod
    def load(cls, directory, *, mmap=True):
        """Read a pack (``shower_basis.py:332-349``)."""
        directory = Path(directory)
        header = json.loads((directory / "pack.json").read_text())
        if header.get("schema") != PACK_SCHEMA:
            raise ValueError("unsupported order >=2 kernel pack")
        store = cls(header["dtype"])
        for item in header["bundles"]:
            meta = dict(item["meta"])
            for name in ("pair_lambda", "pair_mu"):
                meta[name] = np.ascontiguousarray(meta[name], np.int64)
            for name in ("omega", "log_r", "absorption", "speed"):
                meta[name] = np.asarray(meta[name], float)
            key = (tuple(item["key"][0]), int(item["key"][1]), tuple(item["key"][2]))
            store._bundles[key] = dict(meta, coefficients=np.load(
                directory / item["file"], mmap_mode="r" if mmap else None))
        return store


def frequency_indices(omega_grid, frequency_nodes):
    """Evenly spread indices of the cache frequencies (``shower_basis.py:352-361``)."""
    if not isinstance(frequency_nodes, int) or isinstance(frequency_nodes, bool):
        raise TypeError("frequency_nodes must be an integer")
    if not 4 <= frequency_nodes <= len(omega_grid):
        raise ValueError("frequency_nodes must lie between 4 and the cache grid size")
    indices = np.unique(np.rint(np.linspace(0, len(omega_grid) - 1,
                                            frequency_nodes)).astype(int))
    if len(indices) != frequency_nodes:
        raise ValueError("frequency_nodes exceed distinct cache frequencies")
    return indices
