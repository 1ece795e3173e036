"""Medium, wavelength quadrature, detector and receiver contracts.

Mirrors the reference's public input types — ``WavelengthQuadrature``,
``SpectralMedium``, ``DetectorArray`` (``model.py:21-171``), ``SphereSurface``
(``receiver.py:6-51``) and the acceptance Legendre coefficients
(``cache.py:81-99``, ``directional.py:107-130``, ``transport.py:496-530``) —
with the same validation and error messages.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Callable

import numpy as np
from numpy.polynomial.legendre import leggauss


def _axis(values, name, *, positive=False):
    array = np.asarray(values, float)
    if array.ndim != 1 or not len(array) or not np.isfinite(array).all():
        raise ValueError(f"{name} must be a finite nonempty vector")
    if positive and np.any(array <= 0):
        raise ValueError(f"{name} must be positive")
    return array


@dataclass(frozen=True)
class WavelengthQuadrature:
    """Nodes and weights in nm (``model.py:21-44``)."""
    wavelength_nm: np.ndarray
    weight_nm: np.ndarray

    def __post_init__(self):
        wavelength = _axis(self.wavelength_nm, "wavelength_nm", positive=True)
        weight = _axis(self.weight_nm, "weight_nm", positive=True)
        if wavelength.shape != weight.shape or np.any(np.diff(wavelength) <= 0):
            raise ValueError("quadrature wavelengths must increase and match weights")
        object.__setattr__(self, "wavelength_nm", wavelength)
        object.__setattr__(self, "weight_nm", weight)

    @classmethod
    def gauss_legendre(cls, low_nm, high_nm, nodes=9):
        if not 0 < low_nm < high_nm:
            raise ValueError("require 0 < low_nm < high_nm")
        if isinstance(nodes, bool) or not isinstance(nodes, (int, np.integer)) or nodes < 1:
            raise ValueError("nodes must be a positive integer")
        x, w = leggauss(int(nodes))
        half = (high_nm - low_nm) / 2
        return cls(low_nm + half * (x + 1), half * w)


@dataclass(frozen=True)
class SpectralMedium:
    """Optical coefficients tabulated on wavelength nodes, linear in between (``model.py:46-91``)."""
    wavelength_nm: np.ndarray
    absorption_per_m: np.ndarray
    scattering_per_m: np.ndarray
    phase_index: np.ndarray
    group_index: np.ndarray
    g: float = 0.9
    provenance: str = "user-specified spectral medium"

    def __post_init__(self):
        wavelength = _axis(self.wavelength_nm, "wavelength_nm", positive=True)
        if len(wavelength) < 2 or np.any(np.diff(wavelength) <= 0):
            raise ValueError("wavelength_nm must contain at least two increasing nodes")
        object.__setattr__(self, "wavelength_nm", wavelength)
        for name in ("absorption_per_m", "scattering_per_m", "phase_index", "group_index"):
            value = _axis(getattr(self, name), name, positive=name != "scattering_per_m")
            if value.shape != wavelength.shape or (name == "scattering_per_m" and np.any(value < 0)):
                raise ValueError(f"{name} must match wavelength_nm")
            object.__setattr__(self, name, value)
        if not np.isfinite(self.g) or not -1 < self.g < 1:
            raise ValueError("HG requires finite -1 < g < 1")

    @property
    def wavelength_range_nm(self):
        return float(self.wavelength_nm[0]), float(self.wavelength_nm[-1])

    def sample(self, wavelength_nm):
        wavelength = np.asarray(wavelength_nm, float)
        low, high = self.wavelength_range_nm
        if not np.isfinite(wavelength).all() or np.any((wavelength < low) | (wavelength > high)):
            raise ValueError(f"wavelength outside medium range [{low:g}, {high:g}] nm")

        def interp(values):
            return np.interp(wavelength, self.wavelength_nm, values)
        return {"absorption_per_m": interp(self.absorption_per_m),
                "scattering_per_m": interp(self.scattering_per_m),
                "phase_index": interp(self.phase_index),
                "group_index": interp(self.group_index)}


@dataclass(frozen=True)
class SphereSurface:
    """Spherical glass of one radius or a radius per OM (``receiver.py:6-26``)."""
    radius_m: np.ndarray | float

    def __post_init__(self):
        radius = np.asarray(self.radius_m, float)
        if radius.ndim > 1 or not np.isfinite(radius).all() or np.any(radius <= 0):
            raise ValueError('surface radius_m must be finite and positive')
        object.__setattr__(self, 'radius_m', radius.copy())

    def radii(self, count):
        return np.broadcast_to(self.radius_m, (count,))

    def subset(self, indices):
        return self if self.radius_m.ndim == 0 else SphereSurface(self.radius_m[indices])


@dataclass(frozen=True)
class DetectorArray:
    """OMs with a common angular shape and spectral response (``model.py:94-171``).

    ``orientations`` point towards the accepted hemisphere; the look direction
    of the transport is ``-orientation``.
    """
    positions_m: np.ndarray
    orientations: np.ndarray
    effective_area_m2: np.ndarray | float
    angular_acceptance: Callable[[np.ndarray], np.ndarray]
    spectral_efficiency: Callable[[np.ndarray], np.ndarray]
    identifiers: dict = field(default_factory=dict)
    provenance: str = "user-specified detector"
    receiver_surface: SphereSurface | None = None

    def __post_init__(self):
        positions = np.asarray(self.positions_m, float)
        orientations = np.asarray(self.orientations, float)
        if (positions.ndim != 2 or positions.shape[1] != 3 or not len(positions)
                or not np.isfinite(positions).all() or orientations.shape != positions.shape
                or not np.isfinite(orientations).all()):
            raise ValueError("positions_m and orientations must be finite (N,3) arrays")
        norm = np.linalg.norm(orientations, axis=1)
        if np.any(norm <= 0):
            raise ValueError("detector orientations must be nonzero")
        orientations = orientations / norm[:, None]
        area = np.broadcast_to(np.asarray(self.effective_area_m2, float), (len(positions),)).copy()
        if not np.isfinite(area).all() or np.any(area <= 0):
            raise ValueError("effective_area_m2 must be finite and positive")
        if self.receiver_surface is not None:
            if not isinstance(self.receiver_surface, SphereSurface):
                raise ValueError('unsupported receiver_surface geometry')
            self.receiver_surface.radii(len(positions))
        for name, value in self.identifiers.items():
            if np.asarray(value).shape != (len(positions),):
                raise ValueError(f"identifier {name!r} must contain one value per OM")
        probe = np.linspace(-1, 1, 17)
        angular = np.asarray(self.angular_acceptance(probe), float)
        spectral = np.asarray(self.spectral_efficiency(np.array([350., 450., 550.])), float)
        if (angular.shape != probe.shape or not np.isfinite(angular).all()
                or np.any(angular < 0) or spectral.shape != (3,)
                or not np.isfinite(spectral).all() or np.any(spectral < 0)):
            raise ValueError("detector response functions must return finite nonnegative arrays")
        object.__setattr__(self, "positions_m", positions)
        object.__setattr__(self, "orientations", orientations)
        object.__setattr__(self, "effective_area_m2", area)

    def __len__(self):
        return len(self.positions_m)

    def subset(self, indices):
        indices = np.asarray(indices)
        return replace(self, positions_m=self.positions_m[indices],
                       orientations=self.orientations[indices],
                       effective_area_m2=self.effective_area_m2[indices],
                       identifiers={k: np.asarray(v)[indices] for k, v in self.identifiers.items()},
                       receiver_surface=(None if self.receiver_surface is None else
                                         self.receiver_surface.subset(indices)))

    def spectral_weight(self, wavelength_nm):
        value = np.asarray(self.spectral_efficiency(np.asarray(wavelength_nm, float)), float)
        if not np.isfinite(value).all() or np.any(value < 0):
            raise ValueError("nonfinite or negative detector spectral response")
        return value

    @property
    def looks(self):
        """Look directions of the transport (``-orientation``)."""
        return np.ascontiguousarray(-self.orientations)


def acceptance_coefficients(acceptance, degree, *, quadrature_order=2048):
    """``alpha_l = 2 pi int A(x) P_l(x) dx`` (``cache.py:81-99``)."""
    from scipy.special import eval_legendre, roots_legendre
    x, w = roots_legendre(int(quadrature_order))
    values = np.asarray(acceptance(x), dtype=float)
    if values.shape != x.shape or not np.isfinite(values).all():
        raise ValueError("acceptance must map an array of cosines to finite values")
    if np.any(values < 0):
        raise ValueError("acceptance must be nonnegative")
    ell = np.arange(int(degree) + 1)
    return 2 * np.pi * (eval_legendre(ell[:, None], x[None, :]) * (w * values)).sum(axis=1)


def acceptance_alpha(detector, *, scattering_degree=8, acceptance_degree=None, tolerance=1e-10):
    """Band-limited acceptance coefficients (``transport.py:496-530``, ``directional.py:107-130``)."""
    probe = max(2 * int(scattering_degree), 64)
    alpha = acceptance_coefficients(detector.angular_acceptance, probe)
    peak = float(np.max(np.abs(alpha)))
    if peak <= 0:
        raise ValueError("acceptance is identically zero")
    significant = np.flatnonzero(np.abs(alpha) > tolerance * peak)
    bandwidth = int(significant[-1]) if len(significant) else 0
    if bandwidth >= probe and acceptance_degree is None:
        raise ValueError("the module acceptance is not band-limited")
    if acceptance_degree is not None:
        bandwidth = int(acceptance_degree)
    if bandwidth > scattering_degree:
        raise ValueError(f"module acceptance needs degree {bandwidth}, above "
                         f"scattering_degree={scattering_degree}")
    return alpha[:bandwidth + 1].copy()
