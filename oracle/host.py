"""Host-side preparations of the route, restated for the oracle — TEST INFRASTRUCTURE ONLY.

The oracle route builds every host input itself instead of taking the
package's: the shower frame, the optical panels of the direct light, the
node-band choice of the band fold, the segment bounds behind time origins
and module selection, and the delayed-light groups. Each function cites the
reference code it follows.

Pins (tests/test_oracle_host.py): the frame's axis is the dominant eigenvector
of the light-weighted covariance (checked on a line source and against the
mean direction of a point source); the panels reproduce the efficiency at
their probes to the stated tolerance and are the medium nodes plus bisections;
node bands at full rank interpolate exactly and the fold residual is at most
one; segment bounds equal a brute-force scan of sample points; the delay
groups of a shifted copy equal those of the original.
"""
from __future__ import annotations

from dataclasses import dataclass
from itertools import combinations

import numpy as np

C_M_PER_NS = 0.299792458


# ---------------------------------------------------------------- the frame

@dataclass(frozen=True)
class Frame:
    """Centre, axis and transverse pair (``AxisFrame``, ``axial_source.py:80-141``)."""
    centre_m: np.ndarray
    axis: np.ndarray
    first: np.ndarray
    second: np.ndarray

    def coordinates(self, field, fraction=0.5):
        """Axial ``z``, transverse offset, direction cosine and time of the point at
        ``fraction`` of every chord (``axial_source.py:126-141``)."""
        length = np.asarray(field.length_m, float)
        point = np.asarray(field.start_m, float) + (fraction * length)[:, None] * field.direction
        offset = point - self.centre_m
        z = offset @ self.axis
        start, end = np.asarray(field.start_ns, float), np.asarray(field.end_ns, float)
        return (z, offset - z[:, None] * self.axis, np.asarray(field.direction) @ self.axis,
                start + fraction * (end - start))

    def rotate(self, vectors):
        """Components along (first, second, axis)."""
        vectors = np.asarray(vectors, float)
        return np.stack([vectors @ self.first, vectors @ self.second, vectors @ self.axis], -1)


def axis_frame(field):
    """``AxisFrame.of`` (``axial_source.py:88-118``): the photon-weighted centre and the
    principal axis of the chord midpoints, pointing along the mean direction; the
    mean (then the dominant angular) direction for a source without spatial extent.
    The transverse pair is axis x helper, helper the z axis unless the axis lies
    within acos(0.9) of it (then x)."""
    midpoint = (np.asarray(field.start_m, float)
                + 0.5 * np.asarray(field.length_m, float)[:, None] * field.direction)
    share = np.asarray(field.photons, float)
    share = share / share.sum()
    centre = share @ midpoint
    offset = midpoint - centre
    covariance = (offset * share[:, None]).T @ offset
    values, vectors = np.linalg.eigh(covariance)
    direction = np.asarray(field.direction, float)
    if np.abs(values).max() <= 1e-24:
        axis = share @ direction
        if np.linalg.norm(axis) <= 1e-12:
            axis = np.linalg.eigh((direction * share[:, None]).T @ direction)[1][:, -1]
    else:
        axis = vectors[:, np.argmax(values)]
    axis = axis / np.linalg.norm(axis)
    if share @ (direction @ axis) < 0:
        axis = -axis
    helper = np.array([1.0, 0.0, 0.0]) if abs(axis[2]) > 0.9 else np.array([0.0, 0.0, 1.0])
    first = np.cross(axis, helper)
    first = first / np.linalg.norm(first)
    return Frame(centre, axis, first, np.cross(axis, first))


# ------------------------------------------------------- optical panels

_NODES = np.array([0.0, 1 / 3, 2 / 3, 1.0])
_PROBES = np.array([1 / 8, 1 / 6, 3 / 8, 0.5, 5 / 8, 5 / 6, 7 / 8])


def _cubic_panel(efficiency, left, right, relative, absolute):
    """Cubic in q = (lambda - left)/(right - left) through four nodes; whether it
    holds at the seven probes and stays nonnegative on [0, 1]."""
    width = right - left
    coefficients = np.linalg.solve(np.vander(_NODES, 4, increasing=True),
                                   np.asarray(efficiency(left + width * _NODES), float))
    exact = np.asarray(efficiency(left + width * _PROBES), float)
    fitted = np.polynomial.polynomial.polyval(_PROBES, coefficients)
    good = np.all(np.abs(fitted - exact) <= absolute + relative * np.abs(exact))
    stationary = np.polynomial.polynomial.polyroots(coefficients[1:] * [1, 2, 3])
    inside = [r.real for r in np.atleast_1d(stationary) if abs(r.imag) < 1e-12 and 0 < r.real < 1]
    low = np.polynomial.polynomial.polyval(np.r_[0.0, inside, 1.0], coefficients).min()
    return coefficients, bool(good and low >= -absolute)


def prepare_optics(medium, detector, wavelength_range_nm=None, *, relative=1e-8, absolute=1e-12,
                   depth=16):
    """Panels of the direct-light integral (``prepare_optics``, ``ballistic_interval.py:
    64-161``, cubic representation): the range ends, the medium nodes and a grid of
    at most 20 nm; a panel whose cubic misses the efficiency at a probe (or dips
    below zero) is halved, at most ``depth`` times."""
    low, high = wavelength_range_nm or (float(medium.wavelength_nm[0]),
                                        float(medium.wavelength_nm[-1]))
    start = np.unique(np.r_[low, high, medium.wavelength_nm,
                            np.linspace(low, high, max(2, int(np.ceil((high - low) / 20)) + 1))])
    start = start[(start >= low) & (start <= high)]
    efficiency = detector.spectral_weight
    panels = []

    def split(left, right, level):
        coefficients, good = _cubic_panel(efficiency, left, right, relative, absolute)
        if good:
            panels.append((left, right, coefficients))
            return
        if level == depth:
            raise ValueError("spectral-efficiency interpolation did not converge")
        middle = left + (right - left) / 2
        split(left, middle, level + 1)
        split(middle, right, level + 1)

    for left, right in zip(start[:-1], start[1:]):
        split(left, right, 0)
    edges = np.array([p[0] for p in panels] + [panels[-1][1]])
    sample = medium.sample(edges)
    return dict(panel_edges=edges, phase_nodes=np.asarray(sample["phase_index"], float),
                group_nodes=np.asarray(sample["group_index"], float),
                extinction_nodes=np.asarray(sample["absorption_per_m"], float)
                + np.asarray(sample["scattering_per_m"], float),
                efficiency_coefficients=np.array([p[2] for p in panels]))


# ------------------------------------------------------------ node bands

@dataclass(frozen=True)
class NodeBands:
    """The fields of ``BandFold`` the fold re-fit and the compile read."""
    bands: tuple
    interpolation: np.ndarray     # (degree+1, wavelength, rank)
    scale: np.ndarray             # weight x efficiency per wavelength
    phase_index: np.ndarray
    worst_design_residual: float
    best_design_residual: float


def node_bands(wavelength_nm, weight_nm, efficiency, phase_index, degree, rank,
               beta_design_min=0.8, bands=None):
    """``BandFold.of`` (``band_fold.py:82-130``) by exhaustive search: the ``rank``
    wavelengths whose columns F_l(lambda; beta) = w eta / lambda^2 (1 - c^2)
    P_l(c), c = 1/(beta n_lambda), best interpolate every other wavelength (least
    squares per degree), judged by the worst relative residual over beta above
    ``beta_design_min`` on a grid dense in 1/beta - 1. The columns are built with
    SciPy's Legendre polynomials, not the reference's recurrence.

    The worst residuals of the best few subsets can agree to ~1e-8, below the
    rounding of the pseudo-inverse of the degree-0 and degree-1 columns (which
    span only two functions of beta): which of them wins depends on the last
    bits of the arithmetic. ``bands`` then fixes the subset (the interpolation
    is still computed here), and ``best_design_residual`` records the minimum
    over all subsets for the optimality check."""
    from scipy.special import eval_legendre
    wavelength = np.asarray(wavelength_nm, float)
    scale = np.asarray(weight_nm, float) * np.asarray(efficiency, float)
    phase = np.asarray(phase_index, float)
    beta = 1.0 / (1.0 + np.r_[0.0, np.geomspace(1e-7, phase.max() - 1.0, 800)])
    cosine = 1.0 / np.outer(phase, beta)                           # (lambda, beta)
    light = (scale / wavelength ** 2)[:, None] * np.clip(1.0 - cosine ** 2, 0.0, None)
    columns = np.stack([eval_legendre(ell, np.clip(cosine, -1.0, 1.0)) * light
                        for ell in range(int(degree) + 1)])      # (l, lambda, beta)
    norm = np.maximum(np.sqrt((light ** 2).sum(axis=0)), 1e-300)
    design = beta > beta_design_min

    def fit(subset):
        # the least-squares interpolation through the pseudo-inverse, as the
        # reference forms it: the worst residuals of the best few subsets agree
        # to ~1e-8, so the choice among them follows this exact arithmetic
        nodes = columns[:, list(subset), :]
        mixing = columns @ np.linalg.pinv(nodes)                    # (l, lambda, rank)
        residual = np.linalg.norm(columns - mixing @ nodes, axis=1) / norm[None, :]
        return mixing, residual

    scores = {subset: float(fit(subset)[1][:, design].max())
              for subset in combinations(range(len(wavelength)), int(rank))}
    best = min(scores, key=scores.get)
    chosen = best if bands is None else tuple(int(b) for b in bands)
    mixing, residual = fit(chosen)
    return NodeBands(tuple(int(b) for b in chosen), mixing, scale, phase,
                     float(residual[:, design].max()), scores[best])


# ---------------------------------------------------------- segment bounds

def _closest_squared(positions, elements):
    """(OM, element) squared distance to the nearest point of each segment."""
    start = np.asarray(elements.start_m, float)
    direction = np.asarray(elements.direction, float)
    length = np.asarray(elements.length_m, float)
    offset = np.asarray(positions, float)[:, None, :] - start[None]
    along = np.clip(np.einsum("oek,ek->oe", offset, direction), 0.0, length[None])
    gap = offset - along[..., None] * direction[None]
    return np.einsum("oek,oek->oe", gap, gap)


def min_distance(positions, elements, block=256):
    """Distance of every OM to the nearest element segment
    (``min_distance_to_segments``, ``_prompt_numba.py:2221-2243``)."""
    positions = np.asarray(positions, float)
    out = np.empty(len(positions))
    for lo in range(0, len(positions), block):
        out[lo:lo + block] = np.sqrt(_closest_squared(positions[lo:lo + block], elements)
                                     .min(axis=1))
    return out


def earliest_arrival(positions, elements, group_index, block=256):
    """Per-OM bound start_ns + n_g/c x distance to the segment, minimised over
    elements (``earliest_element_arrival``, ``_shower_basis.py:51-78``)."""
    positions = np.asarray(positions, float)
    flight = float(group_index) / C_M_PER_NS
    start_ns = np.asarray(elements.start_ns, float)
    out = np.empty(len(positions))
    for lo in range(0, len(positions), block):
        distance = np.sqrt(_closest_squared(positions[lo:lo + block], elements))
        out[lo:lo + block] = (start_ns[None] + flight * distance).min(axis=1)
    return out


# ------------------------------------------------------------ delayed light

def delay_groups(elements, *, prompt_delay_ns, cell_ns, window_end_ns, group_index_max):
    """``delay_groups`` (``shower/ge2.py:122-147``): the delay of a step behind the
    light front of the earliest step; prompt steps (delay <= prompt_delay_ns),
    delayed steps in ``cell_ns`` cells with the cell's smallest delay as clock
    shift, and steps whose light cannot reach the window (beyond)."""
    start = np.asarray(elements.start_m, float)
    t0 = np.asarray(elements.start_ns, float)
    first = int(np.argmin(t0))
    distance = np.sqrt(((start - start[first]) ** 2).sum(axis=1))
    delay = t0 - t0[first] - distance / C_M_PER_NS
    prompt = delay <= prompt_delay_ns
    beyond = ~prompt & (delay - (group_index_max - 1.0) * distance / C_M_PER_NS > window_end_ns)
    within = ~prompt & ~beyond
    cells = np.floor(delay / cell_ns).astype(np.int64)
    groups = [(within & (cells == cell), float(delay[within & (cells == cell)].min()))
              for cell in np.unique(cells[within])]
    return prompt, groups, beyond


@dataclass(frozen=True)
class AxialSource:
    """Channels (points, channels, frequency) of one band on frame points
    (``AxialSource``, ``axial_source.py:144-200``)."""
    frame: Frame
    z_m: np.ndarray
    transverse_m: np.ndarray
    channels: np.ndarray
    degree: int
    azimuthal_degree: int
    reference_ns: float
    summary: dict

    def points_m(self):
        f = self.frame
        return (f.centre_m[None, :] + self.z_m[:, None] * f.axis[None, :]
                + self.transverse_m[:, :1] * f.first[None, :]
                + self.transverse_m[:, 1:2] * f.second[None, :])
