"""Synthetic media, arrays and the benchmark workloads.

``synthetic_medium_model`` and ``baikal_like_detector`` mirror the reference's
public demonstration inputs (``examples/synthetic.py:12-54``). The
detector-scale array and the workload builders freeze the choices the
workload descriptions leave open (cluster layout, vertex rule):

* ``detector_scale_array(clusters)``: clusters on a line along x every 300 m;
  each cluster has 7 strings on a 60 m-radius heptagon plus a centre string,
  36 OMs per string every 15 m, looking down like the demonstration array.
* ``plain_small()``: 48 OMs, 2000-element shower of seed 7 at the array
  centroid + (-12, 8, 0) m along (0.35, -0.15, 0.925), 300 m charged track;
  5 ns bins over -60..745 ns; 3 wavelength nodes.
* ``plain_detector_scale()``: 2304 OMs, 50,000 elements, 1500 m charged track,
  vertex uniform in the OM bounding box and isotropic direction drawn from
  ``default_rng(11)`` (vertex first, then direction); 1 ns bins; 8 nodes.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .model import DetectorArray, SpectralMedium, SphereSurface, WavelengthQuadrature
from .sources import SourcePose, SyntheticShower


def synthetic_medium_model():
    return SpectralMedium(
        [400., 450., 500.], [.030, .020, .040], [.030, .022, .017],
        [1.344, 1.339, 1.336], [1.386, 1.374, 1.367], g=.9,
        provenance="synthetic Baikal-like demo; not a calibration")


def angular_acceptance(cosine):
    x = np.asarray(cosine, float)
    return .45 + .45 * x + .10 * x * x


def spectral_efficiency(wavelength_nm):
    wavelength = np.asarray(wavelength_nm, float)
    return .20 * np.exp(-0.5 * ((wavelength - 450.) / 90.) ** 2)


def _array(positions, identifiers, provenance):
    count = len(positions)
    return DetectorArray(positions, np.tile([0., 0., -1.], (count, 1)), np.pi * .20 ** 2,
                         angular_acceptance, spectral_efficiency,
                         receiver_surface=SphereSurface(.20), identifiers=identifiers,
                         provenance=provenance)


def baikal_like_detector(*, clusters=1, strings=8, modules=6, cluster_pitch_m=180.,
                         string_radius_m=25., vertical_spacing_m=15.):
    """The demonstration array (``examples/synthetic.py:19-54``)."""
    cluster_centres = np.zeros((clusters, 3))
    cluster_centres[:, 0] = np.arange(clusters) * cluster_pitch_m
    azimuth = np.arange(strings) * 2 * np.pi / strings
    string_offsets = np.column_stack((string_radius_m * np.cos(azimuth),
                                      string_radius_m * np.sin(azimuth), np.zeros(strings)))
    heights = (np.arange(modules) - (modules - 1) / 2) * vertical_spacing_m
    positions = (cluster_centres[:, None, None, :] + string_offsets[None, :, None, :]
                 + np.column_stack((np.zeros(modules), np.zeros(modules), heights))[None, None, :, :]
                 ).reshape(-1, 3)
    identifiers = {
        "cluster_id": np.repeat(np.arange(clusters), strings * modules),
        "string_id": np.tile(np.repeat(np.arange(strings), modules), clusters),
        "module_id": np.tile(np.arange(modules), clusters * strings),
    }
    return _array(positions, identifiers, "synthetic Baikal-like geometry and OM; not a calibration")


def detector_scale_array(clusters=8, *, modules=36, cluster_pitch_m=300., radius_m=60.,
                         vertical_spacing_m=15.):
    """Clusters of a heptagon of strings plus a centre string (8 strings x ``modules``)."""
    azimuth = np.arange(7) * 2 * np.pi / 7
    offsets = np.vstack([np.column_stack((radius_m * np.cos(azimuth), radius_m * np.sin(azimuth),
                                          np.zeros(7))), np.zeros((1, 3))])
    heights = (np.arange(modules) - (modules - 1) / 2) * vertical_spacing_m
    centres = np.zeros((clusters, 3))
    centres[:, 0] = (np.arange(clusters) - (clusters - 1) / 2) * cluster_pitch_m
    positions = (centres[:, None, None, :] + offsets[None, :, None, :]
                 + np.column_stack((np.zeros(modules), np.zeros(modules), heights))[None, None, :, :]
                 ).reshape(-1, 3)
    identifiers = {
        "cluster_id": np.repeat(np.arange(clusters), 8 * modules),
        "string_id": np.tile(np.repeat(np.arange(8), modules), clusters),
        "module_id": np.tile(np.arange(modules), clusters * 8),
    }
    return _array(positions, identifiers, "synthetic detector-scale array; not a calibration")


@dataclass(frozen=True)
class Workload:
    """Inputs of one benchmark event."""
    name: str
    medium: SpectralMedium
    detector: DetectorArray
    wavelength: WavelengthQuadrature
    edges_ns: np.ndarray
    omega_per_ns: np.ndarray
    readout_omega_per_ns: np.ndarray
    elements: object
    source_degree: int = 8
    azimuthal_degree: int = 2
    radial_range_m: tuple = (2.0, 150.0)
    radial_nodes: int = 48
    threshold_pe: float = 0.01


def _shower(pose, track_m, elements, seed):
    from dataclasses import replace
    shower = SyntheticShower.gaussian(pose, charged_track_length_m=track_m, elements=elements,
                                      seed=seed)
    return replace(shower.elements, cone_model="spectral")


def plain_small():
    medium = synthetic_medium_model()
    detector = baikal_like_detector(clusters=1)
    centre = detector.positions_m.mean(axis=0)
    pose = SourcePose(centre + np.array([-12.0, 8.0, 0.0]), [0.35, -0.15, 0.925])
    return Workload("plain-small: 48 OMs, 2000-element shower, 5 ns bins", medium, detector,
                    WavelengthQuadrature.gauss_legendre(400., 500., 3),
                    np.arange(-60.0, 745.0, 5.0), np.linspace(0., 1.2, 65),
                    np.linspace(0., 1.2, 321), _shower(pose, 300.0, 2000, 7))


def plain_detector_scale(elements=50_000):
    medium = synthetic_medium_model()
    detector = detector_scale_array(8)
    rng = np.random.default_rng(11)
    low, high = detector.positions_m.min(axis=0), detector.positions_m.max(axis=0)
    vertex = low + rng.uniform(size=3) * (high - low)
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    pose = SourcePose(vertex, direction)
    return Workload(f"plain-detector-scale: 2304 OMs, {elements}-element shower, 1 ns bins",
                    medium, detector, WavelengthQuadrature.gauss_legendre(400., 500., 8),
                    np.arange(-60.0, 745.0, 1.0), np.linspace(0., 1.2, 65),
                    np.linspace(0., 1.2, 321), _shower(pose, 1500.0, elements, 11))


def heavy_event(elements=1_000_000):
    """The heavy single event: 16 wavelength nodes, 1 ns bins over -60..1500 ns,
    a 16-cluster array (4608 OMs), one Gaussian shower of ``elements`` elements
    with a 10 km charged track, vertex and direction from seed 31 as for the
    detector-scale event."""
    medium = synthetic_medium_model()
    detector = detector_scale_array(16)
    rng = np.random.default_rng(31)
    low, high = detector.positions_m.min(axis=0), detector.positions_m.max(axis=0)
    vertex = low + rng.uniform(size=3) * (high - low)
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    pose = SourcePose(vertex, direction)
    return Workload(f"heavy: 4608 OMs, {elements}-element shower, 16 wavelengths, 1 ns bins",
                    medium, detector, WavelengthQuadrature.gauss_legendre(400., 500., 16),
                    np.arange(-60.0, 1501.0, 1.0), np.linspace(0., 1.2, 65),
                    np.linspace(0., 1.2, 321), _shower(pose, 10_000.0, elements, 31))


def serving_plan(detector, count=4096, *, seed=23, low=500, high=200_000, inflate_m=100.0):
    """The draws of the serving stream without building its showers: per event
    ``(elements, vertex, direction, shower_seed)``. Element counts are
    log-uniform in [low, high], vertices uniform in the array's bounding box
    inflated by ``inflate_m``, directions isotropic, all from ``seed``."""
    rng = np.random.default_rng(seed)
    lo = detector.positions_m.min(axis=0) - inflate_m
    hi = detector.positions_m.max(axis=0) + inflate_m
    plan = []
    for _ in range(count):
        n = int(round(np.exp(rng.uniform(np.log(low), np.log(high)))))
        vertex = lo + rng.uniform(size=3) * (hi - lo)
        direction = rng.normal(size=3)
        direction /= np.linalg.norm(direction)
        plan.append((n, vertex, direction, int(rng.integers(2 ** 31))))
    return plan


def serving_event(entry, track_m=300.0):
    """The shower of one :func:`serving_plan` entry (a ``track_m`` charged track:
    the element count sets the work, not the brightness)."""
    n, vertex, direction, seed = entry
    return _shower(SourcePose(vertex, direction), track_m, n, seed)


def serving_stream(detector, count=4096, *, seed=23, low=500, high=200_000, inflate_m=100.0,
                   track_m=300.0):
    """Events of the serving stream (:func:`serving_plan`), built one at a time:
    yields ``count`` host element sets in order."""
    for entry in serving_plan(detector, count, seed=seed, low=low, high=high,
                              inflate_m=inflate_m):
        yield serving_event(entry, track_m)


def poisson_arrivals(rate_per_s, count, seed=23):
    """Arrival times (s from the first) of a Poisson stream of ``count`` events."""
    gaps = np.random.default_rng(seed).exponential(1.0 / rate_per_s, count)
    gaps[0] = 0.0
    return np.cumsum(gaps)


def synthetic_pack(medium, quadrature, *, degree=8, acceptance_degree=2,
                   omega=np.linspace(0.0, 1.2, 65), radial_range_m=(2.0, 150.0),
                   radial_nodes=48):
    """An order >=2 kernel pack of the reference's layout with a smooth, positive,
    made-up radial dependence (the RTE table build is not part of this package).

    ``K_(l, pair)(r, w) = 0.3^l / (1 + pair) * 0.5 (mu_s r)^2 / (1 + mu_s r)
    * e^{-i w 0.05 r} / (1 + (0.02 r w)^2)`` per wavelength band, written as cubics
    in ``log r`` on ``radial_nodes - 1`` intervals with flight phase.
    Returns ``(wavelengths, indices, bundle)`` for :meth:`KernelPack.add`.
    """
    wavelengths = np.asarray(quadrature.wavelength_nm, float)
    sample = medium.sample(wavelengths)
    mus = np.asarray(sample["scattering_per_m"], float)
    pl, pm = [], []
    for a in range(acceptance_degree + 1):
        for m in range(a + 1):
            pl.append(a)
            pm.append(m)
    log_r = np.linspace(np.log(radial_range_m[0]), np.log(radial_range_m[1]), radial_nodes)
    omega = np.asarray(omega, float)
    nodes = np.array([0.0, 1 / 3, 2 / 3, 1.0])
    nint = radial_nodes - 1
    coefficients = np.zeros((len(wavelengths), nint, degree + 1, len(pl), 4, len(omega)),
                            np.complex64)
    amplitude = np.array([[0.3 ** ell / (1.0 + p) for p in range(len(pl))]
                          for ell in range(degree + 1)])
    for j in range(nint):
        width = log_r[j + 1] - log_r[j]
        t = nodes * width
        radius = np.exp(log_r[j] + t)
        inverse = np.linalg.inv(np.vander(t, 4))
        for band in range(len(wavelengths)):
            g = 0.5 * (mus[band] * radius) ** 2 / (1.0 + mus[band] * radius)
            shape = (g[:, None] * np.exp(-1j * omega[None, :] * 0.05 * radius[:, None])
                     / (1.0 + (0.02 * radius[:, None] * omega[None, :]) ** 2))   # (node, w)
            cubic = inverse @ shape                                              # (4, w)
            coefficients[band, j] = (amplitude[:, :, None, None] * cubic[None, None]).astype(
                np.complex64)
    bundle = dict(degree=int(degree), acceptance_degree=int(acceptance_degree),
                  pair_lambda=np.asarray(pl, np.int64), pair_mu=np.asarray(pm, np.int64),
                  omega=omega, phase="flight", coefficients=coefficients, log_r=log_r,
                  absorption=np.asarray(sample["absorption_per_m"], float),
                  speed=0.299792458 / np.asarray(sample["group_index"], float))
    return wavelengths, np.arange(len(omega)), bundle


def rte_pack(kernel, cache_directory=None):
    """The order >=2 pack of ``kernel`` built from the RTE tables
    (:func:`lighthit_amd.tables.build_pack`), kept on disk in the pack format
    (``KernelPack.save``) under ``cache_directory`` (default: the
    ``LIGHTHIT_AMD_PACK_CACHE`` environment variable, else ``lighthit_amd-packs``
    in the temporary directory) and keyed by a digest of everything the build
    reads, so later processes load it instead of building it again."""
    import hashlib
    import json
    import os
    import tempfile
    from pathlib import Path

    from .fold import KernelPack
    from .tables import build_pack
    c = kernel.config
    wavelengths = np.asarray(kernel.wavelength.wavelength_nm, float)
    samples = [kernel.medium.sample(float(w)) for w in wavelengths]
    described = {
        "wavelength_nm": wavelengths.tolist(),
        "medium": [{k: float(np.asarray(v)) for k, v in sample.items()} for sample in samples],
        "g": float(kernel.medium.g),
        "acceptance_degree": int(kernel.acceptance()["degree"]),
        "omega_per_ns": np.asarray(c.omega_per_ns, float).tolist(),
        "radial_range_m": [float(x) for x in c.radial_range_m],
        "radial_nodes": int(c.radial_nodes), "scattering_degree": int(c.scattering_degree),
        "source_degree": int(c.source_degree), "k_max_per_m": float(c.k_max_per_m),
        "k_panel_per_m": float(c.k_panel_per_m), "k_order": int(c.k_order)}
    digest = hashlib.sha256(json.dumps(described, sort_keys=True).encode()).hexdigest()[:20]
    root = Path(cache_directory or os.environ.get("LIGHTHIT_AMD_PACK_CACHE")
                or Path(tempfile.gettempdir()) / "lighthit_amd-packs")
    folder = root / f"ge2-pack-{digest}"
    if (folder / "pack.json").exists():
        return KernelPack.load(folder)
    pack = KernelPack()
    pack.add(*build_pack(kernel))
    root.mkdir(parents=True, exist_ok=True)
    partial = Path(tempfile.mkdtemp(prefix=folder.name + ".", dir=root))
    pack.save(partial)
    try:
        os.replace(partial, folder)          # another process may have won the race
    except OSError:
        pass
    return KernelPack.load(folder) if (folder / "pack.json").exists() else pack


def workload_route(name, device="cuda", pack="synthetic"):
    """Workload ``name`` ("detector", "small" or "heavy"), its route kernel and the
    route (tables resident on ``device`` on first use): the configuration the
    benchmark and the parity fixtures run. ``pack`` is "synthetic" (the seeded
    made-up pack of :func:`synthetic_pack`, whose cost does not depend on its
    values) or "rte" (the pack built from the RTE tables, :func:`rte_pack`).
    The order-1 selection radius is 300 m (1000 m for the heavy event's 10 km
    track): the default guard of the radial range (150 m) raises for these
    events."""
    from .fold import KernelPack
    from .route import RouteConfig, RouteKernel, ShowerConfig, ShowerPrompt, ShowerTransport
    workload = {"detector": plain_detector_scale, "small": plain_small,
                "heavy": heavy_event}[name]()
    heavy = name == "heavy"
    degree, azimuthal = (16, 4) if heavy else (8, 2)
    kernel = RouteKernel(workload.medium, workload.detector,
                         RouteConfig(relative_time_edges_ns=workload.edges_ns,
                                     wavelength_nodes=len(workload.wavelength.wavelength_nm),
                                     omega_per_ns=workload.omega_per_ns,
                                     readout_omega_per_ns=workload.readout_omega_per_ns,
                                     scattering_degree=degree, source_degree=degree,
                                     azimuthal_degree=azimuthal))
    if pack == "rte":
        tables = rte_pack(kernel)
    elif pack == "synthetic":
        tables = KernelPack()
        tables.add(*synthetic_pack(workload.medium, kernel.wavelength, degree=degree,
                                   omega=workload.omega_per_ns))
    else:
        raise ValueError(f"pack must be 'synthetic' or 'rte', not {pack!r}")
    config = ShowerConfig(prompt=ShowerPrompt(max_distance_m=1000.0 if heavy else 300.0))
    return workload, kernel, ShowerTransport(kernel, tables, config=config, device=device)
