"""Objects with only the reference's attribute sets (transport.py:28-140 KernelConfig,
TransportKernel; model.py:94-171 DetectorArray; receiver.py:6-26 SphereSurface;
model.py:46-92 SpectralMedium): what a reference user hands to the route."""
from dataclasses import dataclass, field, replace

import numpy as np


@dataclass(frozen=True)
class ReferenceKernelConfig:
    omega_per_ns: np.ndarray = field(default_factory=lambda: np.linspace(0.0, 1.2, 321))
    relative_time_edges_ns: np.ndarray = field(
        default_factory=lambda: np.arange(-60.0, 740.0 + 5.0, 5.0))
    wavelength_nodes: int = 32
    wavelength_range_nm: tuple | None = None
    scattering_degree: int = 24
    source_degree: int = 32
    azimuthal_degree: int = 4
    cell_m: float = 0.12
    k_max_per_m: float = 8.0
    k_panel_per_m: float = 0.04
    k_order: int = 10
    radial_range_m: tuple = (3.0, 300.0)
    radial_nodes: int = 220
    angular_backend: str = "auto"
    acceptance_degree: int | None = None
    acceptance_tolerance: float = 1e-10
    receiver_block: int = 8
    threshold_pe: float = 0.01
    detail_threshold_pe: float | None = None
    cache_directory: object = None
    cache_policy: str = "build"
    spectral_folded_cache: object = "auto"
    allow_experimental: bool = False
    readout_omega_per_ns: np.ndarray | None = None
    scattered_receiver: str = "centre"
    receiver_vertex: object = None
    first_entry_tables: object = None


@dataclass(frozen=True)
class ReferenceSphere:
    radius_m: object

    def radii(self, count):
        return np.broadcast_to(np.asarray(self.radius_m, float), (count,))

    def subset(self, indices):
        radius = np.asarray(self.radius_m, float)
        return self if radius.ndim == 0 else ReferenceSphere(radius[indices])


@dataclass(frozen=True)
class ReferenceDetector:
    positions_m: np.ndarray
    orientations: np.ndarray
    effective_area_m2: np.ndarray
    angular_acceptance: object
    spectral_efficiency: object
    identifiers: dict = field(default_factory=dict)
    provenance: str = "user-specified detector"
    receiver_surface: object = None

    def __len__(self):
        return len(self.positions_m)

    def subset(self, indices):
        indices = np.asarray(indices)
        return replace(self, positions_m=self.positions_m[indices],
                       orientations=self.orientations[indices],
                       effective_area_m2=self.effective_area_m2[indices],
                       identifiers={k: np.asarray(v)[indices] for k, v in self.identifiers.items()},
                       receiver_surface=(None if self.receiver_surface is None
                                         else self.receiver_surface.subset(indices)))

    def spectral_weight(self, wavelength_nm):
        return np.asarray(self.spectral_efficiency(np.asarray(wavelength_nm, float)), float)


class ReferenceKernel:
    """medium, detector, config, wavelength and acceptance(), as TransportKernel has."""

    def __init__(self, medium, detector, config, wavelength, alpha):
        self.medium, self.detector, self.config, self.wavelength = (medium, detector, config,
                                                                    wavelength)
        self._alpha = np.asarray(alpha, float)

    def acceptance(self):
        return {"degree": len(self._alpha) - 1, "alpha": self._alpha.copy(),
                "residual_above_degree": 0.0, "full_alpha": self._alpha.copy()}


def reference_shaped_demonstration():
    """The demonstration kernel as a reference user builds it: README.md:76-83 config
    (without readout grid: the route forces it), detector with orientations only."""
    from lighthit_amd.model import WavelengthQuadrature, acceptance_alpha
    from lighthit_amd.synthetic import angular_acceptance, plain_small, spectral_efficiency
    w = plain_small()
    d = w.detector
    detector = ReferenceDetector(np.array(d.positions_m), np.array(d.orientations),
                                 np.array(d.effective_area_m2), angular_acceptance,
                                 spectral_efficiency, receiver_surface=ReferenceSphere(0.2))
    config = ReferenceKernelConfig(
        omega_per_ns=np.linspace(0.0, 1.2, 65), relative_time_edges_ns=w.edges_ns,
        wavelength_nodes=3, scattering_degree=8, source_degree=8, azimuthal_degree=2,
        radial_range_m=(2.0, 150.0), radial_nodes=48, k_max_per_m=4.0, k_panel_per_m=0.1,
        k_order=6, threshold_pe=0.01)
    quadrature = WavelengthQuadrature.gauss_legendre(400.0, 500.0, 3)
    kernel = ReferenceKernel(w.medium, detector, config, quadrature,
                             acceptance_alpha(d, scattering_degree=8))
    return w, kernel
