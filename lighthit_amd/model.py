"""Inputs of the route: medium, wavelength quadrature, OMs and their glass.

The route reads these objects only through the reference's attribute names
(``SpectralMedium``, ``WavelengthQuadrature``, ``DetectorArray`` of
``model.py:21-171``; ``SphereSurface`` of ``receiver.py:6-26``), so the
reference's own objects can be handed to it unchanged. Two views are all the
route takes of a detector beyond those attributes: :func:`look_directions`
(``-orientations``, the photon direction an OM sees head-on, as the
reference's routes form it at ``shower/api.py:271`` and ``prompt.py:530``) and
:func:`glass_radii`.

The classes below are small stand-ins with the same attribute set, for code
without a reference installation (tests, benchmarks, synthetic workloads).
They check what the route relies on (shapes, finiteness, signs) and nothing
else.
"""
from __future__ import annotations

import numpy as np

_MEDIUM_COLUMNS = ("absorption_per_m", "scattering_per_m", "phase_index", "group_index")


def look_directions(detector):
    """Unit photon directions seen head-on by each OM: ``-orientations``, (N, 3)."""
    orientation = np.asarray(detector.orientations, float)
    norm = np.linalg.norm(orientation, axis=1, keepdims=True)
    if np.any(norm <= 0):
        raise ValueError("detector orientations must be nonzero")
    return np.ascontiguousarray(-orientation / norm)


def glass_radii(detector):
    """Radius of every OM's glass sphere, or ``None`` when the detector has no surface."""
    surface = getattr(detector, "receiver_surface", None)
    if surface is None:
        return None
    return np.ascontiguousarray(surface.radii(len(detector.positions_m)), float)


def _finite(values, what, *, shape=None, positive=False, nonnegative=False):
    array = np.array(values, dtype=float)
    if shape is not None and array.shape != shape:
        raise ValueError(f"{what}: shape {array.shape}, expected {shape}")
    if array.size == 0 or not np.all(np.isfinite(array)):
        raise ValueError(f"{what} must be finite and nonempty")
    if positive and not np.all(array > 0):
        raise ValueError(f"{what} must be positive")
    if nonnegative and not np.all(array >= 0):
        raise ValueError(f"{what} must be nonnegative")
    return array


class WavelengthQuadrature:
    """Wavelength nodes and weights, both in nm."""

    def __init__(self, wavelength_nm, weight_nm):
        nodes = _finite(wavelength_nm, "wavelength_nm", positive=True).ravel()
        weights = _finite(weight_nm, "weight_nm", shape=nodes.shape, positive=True)
        if len(nodes) > 1 and not np.all(nodes[1:] > nodes[:-1]):
            raise ValueError("wavelength nodes must increase")
        self.wavelength_nm, self.weight_nm = nodes, weights

    @classmethod
    def gauss_legendre(cls, low_nm, high_nm, nodes=9):
        """Gauss-Legendre rule of ``nodes`` points on [low, high]. The nodes are
        formed as low + half (x + 1) so that they equal the reference's to the
        last bit (the band fold picks node bands by index)."""
        count = int(nodes)
        if count != nodes or count < 1 or isinstance(nodes, bool):
            raise ValueError("nodes must be a positive integer")
        if not 0 < low_nm < high_nm:
            raise ValueError("require 0 < low_nm < high_nm")
        x, w = np.polynomial.legendre.leggauss(count)
        half = (high_nm - low_nm) / 2
        return cls(low_nm + half * (x + 1), half * w)


class SpectralMedium:
    """Water tabulated on wavelength nodes (absorption, scattering, phase and
    group index), linear between nodes, with the HG asymmetry ``g``."""

    def __init__(self, wavelength_nm, absorption_per_m, scattering_per_m, phase_index,
                 group_index, g=0.9, provenance="user-specified spectral medium"):
        nodes = _finite(wavelength_nm, "wavelength_nm", positive=True).ravel()
        if len(nodes) < 2 or not np.all(np.diff(nodes) > 0):
            raise ValueError("wavelength_nm needs two or more increasing nodes")
        table = np.vstack([_finite(values, name, shape=nodes.shape,
                                   nonnegative=name == "scattering_per_m",
                                   positive=name != "scattering_per_m")
                           for name, values in zip(_MEDIUM_COLUMNS, (
                               absorption_per_m, scattering_per_m, phase_index, group_index))])
        if not (np.isfinite(g) and abs(g) < 1):
            raise ValueError("the HG asymmetry g must lie in (-1, 1)")
        self.wavelength_nm = nodes
        self._table = table
        self.absorption_per_m, self.scattering_per_m, self.phase_index, self.group_index = table
        self.g = float(g)
        self.provenance = provenance

    @property
    def wavelength_range_nm(self):
        return float(self.wavelength_nm[0]), float(self.wavelength_nm[-1])

    def sample(self, wavelength_nm):
        """Each column linearly interpolated at ``wavelength_nm`` (inside the table)."""
        at = np.asarray(wavelength_nm, float)
        low, high = self.wavelength_range_nm
        if not np.all(np.isfinite(at)) or np.any(at < low) or np.any(at > high):
            raise ValueError(f"wavelength outside medium range [{low:g}, {high:g}] nm")
        return {name: np.interp(at, self.wavelength_nm, row)
                for name, row in zip(_MEDIUM_COLUMNS, self._table)}


class SphereSurface:
    """Glass sphere of every OM: one radius, or one per OM."""

    def __init__(self, radius_m):
        radius = _finite(radius_m, "surface radius_m", positive=True)
        if radius.ndim > 1:
            raise ValueError("surface radius_m must be a scalar or one value per OM")
        self.radius_m = radius

    def radii(self, count):
        return np.broadcast_to(self.radius_m, (count,))

    def subset(self, indices):
        if self.radius_m.ndim == 0:
            return self
        return SphereSurface(self.radius_m[np.asarray(indices)])


class DetectorArray:
    """OMs: centres, orientations (towards the accepted hemisphere), effective
    areas, a common angular and spectral response and optional glass spheres."""

    def __init__(self, positions_m, orientations, effective_area_m2, angular_acceptance,
                 spectral_efficiency, identifiers=None, provenance="user-specified detector",
                 receiver_surface=None):
        positions = _finite(positions_m, "positions_m")
        if positions.ndim != 2 or positions.shape[1] != 3:
            raise ValueError("positions_m must be an (N, 3) array")
        count = len(positions)
        axis = _finite(orientations, "orientations", shape=positions.shape)
        length = np.linalg.norm(axis, axis=1, keepdims=True)
        if np.any(length <= 0):
            raise ValueError("detector orientations must be nonzero")
        area = _finite(np.broadcast_to(np.asarray(effective_area_m2, float), (count,)),
                       "effective_area_m2", positive=True)
        labels = {key: np.asarray(value) for key, value in (identifiers or {}).items()}
        if any(value.shape != (count,) for value in labels.values()):
            raise ValueError("identifiers need one value per OM")
        if receiver_surface is not None:
            receiver_surface.radii(count)
        self._set(positions, axis / length, area, angular_acceptance, spectral_efficiency,
                  labels, provenance, receiver_surface)

    def _set(self, positions, orientations, area, angular, spectral, labels, provenance,
             surface):
        self.positions_m, self.orientations, self.effective_area_m2 = positions, orientations, area
        self.angular_acceptance, self.spectral_efficiency = angular, spectral
        self.identifiers, self.provenance, self.receiver_surface = labels, provenance, surface

    def __len__(self):
        return len(self.positions_m)

    def subset(self, indices):
        """The OMs ``indices`` with their geometry, areas, labels and glass."""
        pick = np.asarray(indices)
        other = DetectorArray.__new__(DetectorArray)
        other._set(self.positions_m[pick], self.orientations[pick], self.effective_area_m2[pick],
                   self.angular_acceptance, self.spectral_efficiency,
                   {key: value[pick] for key, value in self.identifiers.items()},
                   self.provenance,
                   None if self.receiver_surface is None else self.receiver_surface.subset(pick))
        return other

    def spectral_weight(self, wavelength_nm):
        """Spectral efficiency at ``wavelength_nm`` (finite and nonnegative)."""
        value = np.asarray(self.spectral_efficiency(np.asarray(wavelength_nm, float)), float)
        if not np.all(np.isfinite(value)) or np.any(value < 0):
            raise ValueError("nonfinite or negative detector spectral response")
        return value


def acceptance_alpha(detector, *, scattering_degree=8, acceptance_degree=None, tolerance=1e-10):
    """Legendre coefficients ``alpha_l = 2 pi int_{-1}^{1} A(x) P_l(x) dx`` of the OM
    acceptance, truncated at its band limit (``TransportKernel.acceptance``,
    ``transport.py:496-530``): the largest degree whose coefficient exceeds
    ``tolerance`` of the largest, probed up to max(2 scattering_degree, 64) on a
    2048-point Gauss rule; ``acceptance_degree`` forces the degree; a band limit
    above ``scattering_degree`` is an error."""
    probe = max(2 * int(scattering_degree), 64)
    x, w = np.polynomial.legendre.leggauss(2048)
    response = np.asarray(detector.angular_acceptance(x), float)
    if response.shape != x.shape or not np.all(np.isfinite(response)):
        raise ValueError("acceptance must map an array of cosines to finite values")
    if np.any(response < 0):
        raise ValueError("acceptance must be nonnegative")
    alpha = 2 * np.pi * (np.polynomial.legendre.legvander(x, probe).T @ (w * response))
    peak = float(np.abs(alpha).max())
    if peak <= 0:
        raise ValueError("acceptance is identically zero")
    above = np.abs(alpha) > tolerance * peak
    limit = int(np.nonzero(above)[0].max())
    if acceptance_degree is not None:
        limit = int(acceptance_degree)
    elif limit >= probe:
        raise ValueError("the module acceptance is not band-limited: its Legendre coefficients "
                         f"are still above the tolerance at degree {probe}")
    if limit > scattering_degree:
        raise ValueError(f"module acceptance needs degree {limit}, above "
                         f"scattering_degree={scattering_degree}")
    return alpha[:limit + 1].copy()


def kernel_alpha(kernel):
    """Acceptance coefficients of a route kernel (``kernel.acceptance()["alpha"]``)."""
    return np.ascontiguousarray(kernel.acceptance()["alpha"], float)
