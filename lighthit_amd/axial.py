"""Shower frame and axial-slice source containers.

Mirrors ``AxisFrame`` and ``AxialSource`` of the reference
(``axial_source.py:80-200``): a frame (centre, axis and a transverse pair) and
the order >=2 source of one wavelength band on a set of points of that frame,
as harmonic channels per point and frequency.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def column_bounds(points):
    """Per-column minimum and maximum of an (N, 3) array (column by column: a
    reduction over axis 0 of an (N, 3) array is about ten times slower)."""
    points = np.asarray(points, float)
    return (np.array([points[:, i].min() for i in range(points.shape[1])]),
            np.array([points[:, i].max() for i in range(points.shape[1])]))


def dot3(vectors, direction):
    """``vectors @ direction`` for 3-vectors, summed in one fixed order (a BLAS
    product may split its work by thread count, and so its last bits)."""
    vectors = np.asarray(vectors, float)
    return (vectors[..., 0] * direction[0] + vectors[..., 1] * direction[1]
            + vectors[..., 2] * direction[2])


@dataclass(frozen=True)
class AxisFrame:
    """Centre, axis and transverse pair of a shower (``axial_source.py:80-141``)."""
    centre_m: np.ndarray
    axis: np.ndarray
    first: np.ndarray
    second: np.ndarray

    @classmethod
    def of(cls, elements):
        """Photon-weighted centre and principal axis of the emission points.

        ``elements`` needs ``midpoints_m``, ``photons`` and ``direction``
        (``axial_source.py:88-118``).
        """
        points = elements.midpoints_m
        weights = np.asarray(elements.photons, float)
        share = weights / weights.sum()
        centre = (share[:, None] * points).sum(axis=0)
        offset = points - centre
        covariance = np.einsum("ij,ik->jk", offset * share[:, None], offset)
        values, vectors = np.linalg.eigh(covariance)
        if np.max(np.abs(values)) <= 1e-24:
            axis = np.sum(share[:, None] * elements.direction, axis=0)
            if np.linalg.norm(axis) <= 1e-12:
                _, angular_vectors = np.linalg.eigh(np.einsum(
                    "i,ij,ik->jk", share, elements.direction, elements.direction))
                axis = angular_vectors[:, -1]
        else:
            axis = vectors[:, int(np.argmax(values))]
        axis = axis / np.linalg.norm(axis)
        if float((share * dot3(elements.direction, axis)).sum()) < 0:
            axis = -axis
        helper = np.array([0.0, 0.0, 1.0])
        if abs(float(axis @ helper)) > 0.9:
            helper = np.array([1.0, 0.0, 0.0])
        first = np.cross(axis, helper)
        first /= np.linalg.norm(first)
        return cls(centre, axis, first, np.cross(axis, first))

    def coordinates(self, elements, fraction=0.5):
        """``z`` along the axis, transverse offset, direction cosine and time of
        the point at ``fraction`` of every chord (``axial_source.py:126-141``)."""
        points = (elements.start_m
                  + (fraction * np.asarray(elements.length_m, float))[:, None] * elements.direction)
        offset = points - self.centre_m
        z = dot3(offset, self.axis)
        transverse = offset - z[:, None] * self.axis
        cosine = dot3(elements.direction, self.axis)
        start = np.asarray(elements.start_ns, float)
        time = start + fraction * (np.asarray(elements.end_ns, float) - start)
        return z, transverse, cosine, time

    def matrix(self):
        """Rows (first, second, axis): global -> frame coordinates."""
        return np.stack((self.first, self.second, self.axis))

    def rotate(self, vectors):
        """``(first, second, axis)`` components, the axis as z."""
        vectors = np.asarray(vectors, float)
        return np.stack((dot3(vectors, self.first), dot3(vectors, self.second),
                         dot3(vectors, self.axis)), axis=-1)


def channel_layout(degree, azimuthal_degree):
    """First channel of every degree: ``starts[l]`` (``_shower_basis.py:1153-1156``)."""
    starts = [0]
    for ell in range(degree + 1):
        starts.append(starts[-1] + 2 * min(ell, azimuthal_degree) + 1)
    return np.asarray(starts, np.int64)


@dataclass(frozen=True)
class AxialSource:
    """Channels ``(points, channels, frequency)`` of one band on frame points.

    Field names follow ``axial_source.py:144-160``; ``z_m``/``transverse_m``
    place the points in ``frame``.
    """
    frame: AxisFrame
    z_m: np.ndarray
    transverse_m: np.ndarray
    channels: np.ndarray
    degree: int
    azimuthal_degree: int
    reference_ns: float
    summary: dict = field(default_factory=dict)

    def points_m(self):
        """Positions of the source points in detector coordinates."""
        return (self.frame.centre_m[None, :]
                + self.z_m[:, None] * self.frame.axis[None, :]
                + self.transverse_m[:, 0:1] * self.frame.first[None, :]
                + self.transverse_m[:, 1:2] * self.frame.second[None, :])

    def local_points(self):
        """Positions in frame coordinates (``transverse``, ``z``)."""
        return np.column_stack((self.transverse_m[:, 0], self.transverse_m[:, 1], self.z_m))
