"""CPU pins of the axial-slice compile oracle (order >=2 source)."""
from dataclasses import replace

import numpy as np
import pytest

from lighthit_amd.axial import AxisFrame
from lighthit_amd.sources import SourcePose, SyntheticShower
from oracle import compile as oracle
from oracle import ge2 as oracle_ge2


def shower(elements=400, length=20.0, seed=3):
    """The reference's compile test shower (``tests/test_shower_basis.py:20-23``)."""
    return SyntheticShower.gaussian(SourcePose([12., 3., 0.], [0.3, 0.2, 0.93], 0.),
                                    charged_track_length_m=length, elements=elements, seed=seed)


def test_harmonics_match_the_directional_oracle():
    rng = np.random.default_rng(1)
    vectors = rng.normal(size=(20, 3))
    vectors[0] = [0.0, 0.0, 1.0]
    vectors[1] = [0.0, 0.0, -2.0]
    degree, max_m = 8, 3
    got = oracle.harmonics(vectors, degree, max_m)
    _, kept, _ = oracle.layout(degree, max_m)
    for vector, row in zip(vectors, got):
        full = oracle_ge2.harmonics(vector / np.linalg.norm(vector), degree)
        np.testing.assert_allclose(row, full[kept], rtol=0, atol=1e-13)


def test_legendre_matches_numpy():
    c = np.linspace(-1, 1, 17)
    got = oracle.legendre(c, 9)
    for ell in range(10):
        coefficients = np.zeros(ell + 1)
        coefficients[ell] = 1
        np.testing.assert_allclose(got[:, ell], np.polynomial.legendre.legval(c, coefficients),
                                   rtol=0, atol=1e-14)


@pytest.mark.parametrize("transverse", ["collapse", "stencil"])
@pytest.mark.parametrize("axial_order", [1, 3])
def test_moments_of_the_basis_equal_those_of_the_chords(transverse, axial_order):
    """``tests/test_shower_basis.py:28-63``: the compiled points carry the chords' moments."""
    field = shower().elements.field(0)
    frame = AxisFrame.of(field)
    source = oracle.compile_axial_slices(field, 4, [0.0], azimuthal_degree=2, cell_m=0.3,
                                         axial_order=axial_order, transverse=transverse,
                                         element_order=2, frame=frame)
    points = frame.rotate(source.points_m() - frame.centre_m)
    weight = source.channels[:, 0, 0].real * np.sqrt(4 * np.pi)
    nodes, node_weights = np.polynomial.legendre.leggauss(2)
    exact = dict.fromkeys(("0", "z", "zz", "zzz", "x", "y", "xx", "yy"), 0.0)
    for fraction, node_weight in zip((nodes + 1) / 2, node_weights / 2):
        z, offset, _, _ = frame.coordinates(field, fraction)
        x, y = offset @ frame.first, offset @ frame.second
        q = field.photons * node_weight
        for name, value in (("0", 1), ("z", z), ("zz", z * z), ("zzz", z ** 3), ("x", x),
                            ("y", y), ("xx", x * x), ("yy", y * y)):
            exact[name] += float(np.sum(q * value))
    got = {"0": weight.sum(), "z": weight @ points[:, 2], "zz": weight @ points[:, 2] ** 2,
           "zzz": weight @ points[:, 2] ** 3, "x": weight @ points[:, 0],
           "y": weight @ points[:, 1], "xx": weight @ points[:, 0] ** 2,
           "yy": weight @ points[:, 1] ** 2}
    total = exact["0"]
    assert got["0"] == pytest.approx(total, rel=1e-12)
    assert got["z"] == pytest.approx(exact["z"], rel=1e-10, abs=1e-10 * total)
    if axial_order == 3:
        for name in ("zz", "zzz"):
            assert got[name] == pytest.approx(exact[name], rel=1e-10, abs=1e-10 * total)
    else:
        # linear weights keep z^2 only to O(h^2): a planted-error check of the pin
        assert abs(got["zz"] - exact["zz"]) > 1e-6 * abs(exact["zz"])
    if transverse == "stencil":
        for name in ("x", "y", "xx", "yy"):
            assert got[name] == pytest.approx(exact[name], rel=1e-10, abs=1e-12 * total)
    else:
        assert np.all(np.abs(points[:, :2]) < 1e-12)


def test_time_lattice_error_is_within_its_bound():
    """``tests/test_shower_basis.py:66-82``: the linear tick lattice against the exact
    ``e^{i omega t}`` of the same axial (cloud-in-cell) weights."""
    field = shower().elements.field(0)
    frame = AxisFrame.of(field)
    omega = np.linspace(0.0, 1.2, 13)
    degree, max_m, cell = 6, 2, 0.3
    for step in (0.05, 0.02):
        basis = oracle.compile_axial_slices(field, degree, omega, azimuthal_degree=max_m,
                                            cell_m=cell, axial_order=1, transverse="collapse",
                                            element_order=2, time_step_ns=step, frame=frame)
        # exact phases, same chord nodes and axial weights
        _, _, ell_of = oracle.layout(degree, max_m)
        angular = (oracle.harmonics(frame.rotate(field.direction), degree, max_m)
                   * oracle.legendre(field.cone_cosine, degree)[:, ell_of])
        z_origin = basis.summary["z_origin_m"]
        slices = np.rint((basis.z_m - z_origin) / cell).astype(np.int64)
        exact = np.zeros_like(basis.channels)
        nodes, node_weights = np.polynomial.legendre.leggauss(2)
        for fraction, node_weight in zip((nodes + 1) / 2, node_weights / 2):
            z, _, _, t = frame.coordinates(field, fraction)
            zl = (z - z_origin) / cell
            zb = np.floor(zl).astype(np.int64)
            phase = np.exp(1j * np.outer(t - basis.reference_ns, omega))
            for shift, w in ((0, 1 - (zl - zb)), (1, zl - zb)):
                at = np.searchsorted(slices, zb + shift)
                values = (w * field.photons * node_weight)[:, None, None] * angular[:, :, None] \
                    * phase[:, None, :]
                np.add.at(exact, at, values)
        total = np.abs(exact).sum(axis=(0, 1))
        error = np.abs(basis.channels - exact).sum(axis=(0, 1))
        assert error[0] <= 1e-12 * total[0]
        assert np.all(error <= (omega * step) ** 2 / 8 * total + 1e-12 * total[0])
        assert np.any(error[1:] > 1e-3 * (omega[1:] * step) ** 2 / 8 * total[1:])


@pytest.mark.parametrize("transverse", ["collapse", "stencil"])
def test_charge_compilation_does_not_need_a_time_lattice(transverse):
    """``tests/test_shower_basis.py:85-100``."""
    field = shower().elements.field(0)
    field = replace(field, start_ns=field.start_ns * 1000, end_ns=field.end_ns * 1000)
    options = dict(azimuthal_degree=2, cell_m=.3, transverse=transverse, stencil_threshold_m=.03)
    zero = oracle.compile_axial_slices(field, 6, [0.], **options)
    full = oracle.compile_axial_slices(field, 6, [0., .03], **options)
    np.testing.assert_allclose(zero.points_m(), full.points_m(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(zero.channels[:, :, 0], full.channels[:, :, 0],
                               rtol=2e-12, atol=1e-10)
    assert zero.summary["time_rows"] == zero.summary["slices"]
    assert full.summary["time_rows"] > zero.summary["time_rows"]


def test_thin_slices_keep_one_point_at_their_centroid():
    field = shower().elements.field(0)
    frame = AxisFrame.of(field)
    wide = oracle.compile_axial_slices(field, 4, [0.0], azimuthal_degree=2, frame=frame,
                                       stencil_threshold_m=0.0)
    thin = oracle.compile_axial_slices(field, 4, [0.0], azimuthal_degree=2, frame=frame,
                                       stencil_threshold_m=10.0)
    assert thin.summary["stencil_slices"] == 0
    assert wide.summary["stencil_slices"] == wide.summary["slices"]
    assert len(thin.z_m) == thin.summary["slices"]
    np.testing.assert_allclose(thin.channels[:, 0, 0].sum(), wide.channels[:, 0, 0].sum(),
                               rtol=1e-12)
