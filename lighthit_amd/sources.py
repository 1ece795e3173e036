"""Charged-step sources: the two-field Cherenkov elements and synthetic showers.

Stand-ins for ``SpectralLightElements``, ``SourcePose`` and ``SyntheticShower``
(``sources.py:12-58, 60-78, 110-262, 396-441``) with the same fields, for
code without a reference installation; the route takes the reference's own
element objects as well. The placement (photon-weighted centroid and
principal axis moved onto a pose) and the seeded generator follow the
reference's steps, so a given seed gives bit-identical elements.
"""
from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

#: Fine-structure constant (the reference's ``g4_source.FINE_STRUCTURE``).
FINE_STRUCTURE = 7.2973525693e-3
C_M_PER_NS = 0.299792458


def _vector3(value, what, *, unit=False):
    """A finite 3-vector (scaled to unit length with ``unit``)."""
    vector = np.array(value, dtype=float)
    if vector.shape != (3,) or not np.all(np.isfinite(vector)):
        raise ValueError(f"{what} must be a finite 3-vector")
    if not unit:
        return vector
    length = np.linalg.norm(vector)
    if not length > 0:
        raise ValueError(f"{what} must be nonzero")
    return vector / length


def _principal_axis(elements):
    points = elements.midpoints_m
    weight = np.asarray(elements.coefficient0, float)
    share = weight / weight.sum()
    centre = np.sum(share[:, None] * points, axis=0)
    offset = points - centre
    covariance = np.einsum("i,ij,ik->jk", share, offset, offset)
    values, vectors = np.linalg.eigh(covariance)
    if float(np.max(np.abs(values))) <= 1e-24:
        axis = np.sum(share[:, None] * elements.direction, axis=0)
    else:
        axis = vectors[:, int(np.argmax(values))]
    if float(np.sum(share * (elements.direction @ axis))) < 0:
        axis = -axis
    return centre, axis / np.linalg.norm(axis)


def _rotation_between(source, target):
    source = _vector3(source, "source direction", unit=True)
    target = _vector3(target, "target direction", unit=True)
    cosine = float(np.clip(source @ target, -1.0, 1.0))
    if np.isclose(cosine, 1.0, atol=1e-14):
        return np.eye(3)
    if np.isclose(cosine, -1.0, atol=1e-14):
        helper = np.array([1.0, 0.0, 0.0])
        if abs(float(source @ helper)) > 0.9:
            helper = np.array([0.0, 1.0, 0.0])
        axis = np.cross(source, helper)
        axis /= np.linalg.norm(axis)
        return 2 * np.outer(axis, axis) - np.eye(3)
    cross = np.cross(source, target)
    skew = np.array([[0.0, -cross[2], cross[1]],
                     [cross[2], 0.0, -cross[0]],
                     [-cross[1], cross[0], 0.0]])
    sine2 = float(cross @ cross)
    return np.eye(3) + skew + skew @ skew * ((1.0 - cosine) / sine2)


@dataclass(frozen=True)
class SourcePose:
    """Target centroid, axis and earliest emission time of a source."""
    position_m: np.ndarray
    direction: np.ndarray
    time_ns: float = 0.0

    def __post_init__(self):
        if not np.isfinite(self.time_ns):
            raise ValueError("time_ns must be finite")
        object.__setattr__(self, "position_m", _vector3(self.position_m, "position_m"))
        object.__setattr__(self, "direction", _vector3(self.direction, "direction", unit=True))


@dataclass(frozen=True)
class LightField:
    """Chords with one photon weight and one cone cosine each (the fields of
    ``LightElements`` that the axial-slice compile reads, ``sources.py:193-205``)."""
    start_m: np.ndarray
    direction: np.ndarray
    length_m: np.ndarray
    photons: np.ndarray
    cone_cosine: np.ndarray
    start_ns: np.ndarray
    end_ns: np.ndarray
    row_index: np.ndarray

    def __len__(self):
        return len(self.length_m)

    @property
    def midpoints_m(self):
        return self.start_m + 0.5 * self.length_m[:, None] * self.direction


_PER_ELEMENT = ("length_m", "coefficient0", "coefficient2", "beta", "start_ns", "end_ns",
                "row_index", "uid")


@dataclass(frozen=True)
class SpectralLightElements:
    """Straight emitting chords with ``dN/dlambda = lambda^-2 (C0 + C2 / n^2)``."""
    start_m: np.ndarray
    direction: np.ndarray
    length_m: np.ndarray
    coefficient0: np.ndarray
    coefficient2: np.ndarray
    beta: np.ndarray
    start_ns: np.ndarray
    end_ns: np.ndarray
    row_index: np.ndarray
    uid: np.ndarray
    reference_phase_index: float = 1.35
    provenance: dict | None = None
    cone_model: str = "frozen"

    def __post_init__(self):
        # what the route relies on: (N, 3) geometry, one finite value per
        # element, positive lengths and beta, C0 >= 0 >= C2
        rows = np.array(self.start_m, dtype=float)
        if rows.ndim != 2 or rows.shape[1:] != (3,) or len(rows) == 0:
            raise ValueError("start_m and direction must be nonempty (N,3) arrays")
        axes = np.array(self.direction, dtype=float)
        if axes.shape != rows.shape:
            raise ValueError("start_m and direction must be nonempty (N,3) arrays")
        columns = {name: np.asarray(getattr(self, name)) for name in _PER_ELEMENT}
        for name, column in columns.items():
            if column.shape != (len(rows),) or not np.all(np.isfinite(column.astype(float))):
                raise ValueError(f"{name} must contain one finite value per element")
        length = np.linalg.norm(axes, axis=1)
        signs_ok = (np.all(columns["length_m"] > 0) and np.all(columns["beta"] > 0)
                    and np.all(columns["coefficient0"] >= 0)
                    and np.all(columns["coefficient2"] <= 0))
        if not (np.all(np.isfinite(rows)) and np.all(length > 0) and signs_ok):
            raise ValueError("invalid spectral element geometry or coefficients")
        if self.cone_model not in ("frozen", "spectral"):
            raise ValueError("cone_model must be frozen or spectral")
        object.__setattr__(self, "start_m", rows)
        object.__setattr__(self, "direction", axes / length[:, None])
        for name, column in columns.items():
            object.__setattr__(self, name, column)

    def __len__(self):
        return len(self.length_m)

    @property
    def midpoints_m(self):
        return self.start_m + 0.5 * self.length_m[:, None] * self.direction

    @property
    def centroid_m(self):
        weight = self.coefficient0 / self.coefficient0.sum()
        return np.sum(weight[:, None] * self.midpoints_m, axis=0)

    def field(self, index, *, phase_index=None):
        """One Cherenkov field of the spectrum as a :class:`LightField`
        (``sources.py:193-205``): index 0 carries ``C0``, index 2 ``-C2``;
        the cone cosine is ``1 / (beta n)`` at ``phase_index``."""
        if index not in (0, 2):
            raise ValueError("field index must be 0 or 2")
        phase = self.reference_phase_index if phase_index is None else float(phase_index)
        photons = self.coefficient0 if index == 0 else -self.coefficient2
        return LightField(self.start_m, self.direction, self.length_m,
                          np.asarray(photons, float), 1.0 / (self.beta * phase),
                          self.start_ns, self.end_ns, self.row_index)

    def subset(self, mask):
        mask = np.asarray(mask, bool)
        if mask.shape != (len(self),) or not np.any(mask):
            raise ValueError("subset mask must retain at least one element")
        return replace(self, start_m=self.start_m[mask], direction=self.direction[mask],
                       length_m=self.length_m[mask], coefficient0=self.coefficient0[mask],
                       coefficient2=self.coefficient2[mask], beta=self.beta[mask],
                       start_ns=self.start_ns[mask], end_ns=self.end_ns[mask],
                       row_index=self.row_index[mask], uid=self.uid[mask],
                       provenance={**(self.provenance or {}), "subset_elements": int(mask.sum())})

    def placed(self, pose: SourcePose):
        """Move the photon-weighted centroid and principal axis onto ``pose``."""
        if not isinstance(pose, SourcePose):
            raise TypeError("pose must be a SourcePose")
        source_centroid, source_axis = _principal_axis(self)
        rotation = _rotation_between(source_axis, pose.direction)
        translation = pose.position_m - rotation @ source_centroid
        delay = float(pose.time_ns - np.min(self.start_ns))
        return replace(self, start_m=self.start_m @ rotation.T + translation,
                       direction=self.direction @ rotation.T,
                       start_ns=self.start_ns + delay, end_ns=self.end_ns + delay,
                       provenance={**(self.provenance or {}), "pose": "placed"})


@dataclass(frozen=True)
class G4Shower:
    """A Geant4 shower event read from the G4 step files (``sources.py:344-393``):
    its spectral elements, the file and the event number."""
    elements: SpectralLightElements
    input_path: str = ""
    event: int = 0

    @property
    def extent_m(self):
        ends = np.vstack((self.elements.start_m, self.elements.start_m
                          + self.elements.length_m[:, None] * self.elements.direction))
        return float(np.linalg.norm(ends.max(axis=0) - ends.min(axis=0)))

    def placed(self, pose: SourcePose):
        return replace(self, elements=self.elements.placed(pose))

    @classmethod
    def from_hdf5(cls, path, event=0, *, reference_phase_index=1.35, max_elements=None):
        """Elements of ``event_<event>/tracks`` (rows ``uid``, ``x_m``, ``y_m``,
        ``z_m``, ``t_ns``, ``beta``, ``step_length_m``): every two consecutive
        rows of one particle (same ``uid``) bound one step, emitted along the
        chord between them with the mean of their beta, the Frank-Tamm yield of
        the step's true path (the later row's ``step_length_m``) and the times
        of the two rows; steps with zero chord or path, or below the Cherenkov
        threshold of ``reference_phase_index``, are dropped; ``max_elements``
        keeps the first steps only. The same steps, in the same order, as the
        reference's reader."""
        from pathlib import Path

        from .hdf5 import read_dataset
        path = Path(path)
        rows = read_dataset(path, f"event_{int(event)}/tracks")
        needed = {"uid", "x_m", "y_m", "z_m", "t_ns", "beta", "step_length_m"}
        if rows.dtype.names is None or not needed.issubset(rows.dtype.names):
            raise ValueError(f"G4 tracks must carry {sorted(needed)}")
        uid = np.asarray(rows["uid"], np.int64)
        later = np.flatnonzero(uid[1:] == uid[:-1]) + 1       # second row of each step
        if max_elements is not None:
            later = later[:int(max_elements)]
        earlier = later - 1
        xyz = np.column_stack([np.asarray(rows[k], float) for k in ("x_m", "y_m", "z_m")])
        chord = xyz[later] - xyz[earlier]
        length = np.linalg.norm(chord, axis=1)
        beta_rows = np.asarray(rows["beta"], float)
        beta = 0.5 * (beta_rows[earlier] + beta_rows[later])
        path_m = np.asarray(rows["step_length_m"], float)[later]
        kept = np.flatnonzero((length > 0) & (path_m > 0)
                              & (beta * reference_phase_index > 1))
        c0 = 2 * np.pi * FINE_STRUCTURE * 1e9 * path_m[kept]
        times = np.asarray(rows["t_ns"], float)
        elements = SpectralLightElements(
            xyz[earlier][kept], chord[kept] / length[kept, None], length[kept], c0,
            -c0 / beta[kept] ** 2, beta[kept], times[earlier][kept], times[later][kept],
            later[kept].astype(np.int64), uid[later][kept], reference_phase_index,
            {"file": str(path), "event": int(event), "rows": int(len(rows)),
             "spectral_model": "two fields S0=lambda^-2, S2=lambda^-2*n_phase^-2"})
        return cls(elements, str(path), int(event))


@dataclass(frozen=True)
class SyntheticShower:
    """Seeded synthetic axial cascade (``sources.py:396-441``)."""
    elements: SpectralLightElements
    seed: int = 0

    @classmethod
    def gaussian(cls, pose: SourcePose, *, charged_track_length_m=100.0,
                 longitudinal_sigma_m=1.2, transverse_sigma_m=0.08,
                 elements=64, beta=0.999, reference_phase_index=1.35, seed=0):
        if elements < 3:
            raise ValueError("synthetic shower needs at least three elements")
        if not np.isfinite(beta) or not 0 < beta <= 1:
            raise ValueError("beta must be finite and in (0, 1]")
        if not np.isfinite(reference_phase_index) or reference_phase_index <= 0:
            raise ValueError("reference_phase_index must be finite and positive")
        for name, value in (("charged_track_length_m", charged_track_length_m),
                            ("longitudinal_sigma_m", longitudinal_sigma_m),
                            ("transverse_sigma_m", transverse_sigma_m)):
            if not np.isfinite(value) or value <= 0:
                raise ValueError(f"{name} must be finite and positive")
        rng = np.random.default_rng(seed)
        z = np.linspace(-3 * longitudinal_sigma_m, 3 * longitudinal_sigma_m, int(elements))
        share = np.exp(-0.5 * (z / longitudinal_sigma_m) ** 2)
        share /= share.sum()
        transverse = rng.normal(scale=transverse_sigma_m, size=(len(z), 2))
        angular = rng.normal(scale=0.025, size=(len(z), 2))
        start = np.column_stack((transverse, z))
        direction = np.column_stack((angular, np.ones(len(z))))
        direction /= np.linalg.norm(direction, axis=1)[:, None]
        segment_length = max(0.01, 6 * longitudinal_sigma_m / len(z))
        length = np.full(len(z), segment_length)
        common = 2 * np.pi * FINE_STRUCTURE * 1e9 * charged_track_length_m * share
        time = (z - z.min()) / (beta * C_M_PER_NS)
        raw = SpectralLightElements(
            start, direction, length, common, -common / beta ** 2,
            np.full(len(z), beta), time, time + length / (beta * C_M_PER_NS),
            np.arange(len(z)), np.zeros(len(z), np.int64), reference_phase_index,
            {"source": "SyntheticShower", "seed": int(seed),
             "charged_track_length_m": float(charged_track_length_m)})
        return cls(raw.placed(pose), int(seed))
