"""GPU parity of the axial-slice compile (lh_compile_layout / lh_compile_channels)
against the FP64 oracle."""
import numpy as np
import pytest

from lighthit_amd.axial import AxisFrame
from lighthit_amd.sources import SourcePose, SyntheticShower
from oracle import compile as oracle

pytestmark = pytest.mark.gpu

# complex64 channels of FP64 sums: per value, relative to the value and to the
# largest channel of its frequency
RELATIVE = 1e-6
FLOOR = 1e-7


def _check(got, expected):
    np.testing.assert_allclose(got.z_m, expected.z_m, rtol=0, atol=1e-12)
    np.testing.assert_allclose(got.transverse_m, expected.transverse_m, rtol=0, atol=1e-12)
    assert got.channels.shape == expected.channels.shape
    scale = np.abs(expected.channels).max(axis=(0, 1), keepdims=True)
    bound = RELATIVE * np.abs(expected.channels) + FLOOR * scale
    excess = np.abs(got.channels - expected.channels) - bound
    assert np.all(excess <= 0), float(excess.max())


def shower(elements=400, length=20.0, seed=3):
    return SyntheticShower.gaussian(SourcePose([12., 3., 0.], [0.3, 0.2, 0.93], 0.),
                                    charged_track_length_m=length, elements=elements, seed=seed)


@pytest.mark.parametrize("transverse", ["collapse", "stencil"])
@pytest.mark.parametrize("axial_order", [1, 3])
@pytest.mark.parametrize("omega", [[0.0], list(np.linspace(0.0, 1.2, 13))])
def test_field_compile_matches_oracle(device, transverse, axial_order, omega):
    from lighthit_amd.compile import compile_axial_slices
    field = shower().elements.field(0)
    frame = AxisFrame.of(field)
    options = dict(azimuthal_degree=2, cell_m=0.3, axial_order=axial_order,
                   transverse=transverse, element_order=2, frame=frame,
                   stencil_threshold_m=0.03)
    expected = oracle.compile_axial_slices(field, 4, omega, **options)
    got = compile_axial_slices(field, 4, omega, **options)
    _check(got, expected)


@pytest.mark.parametrize("time_deposit", ["linear", "nearest"])
def test_route_basis_on_the_demonstration_shower(device, time_deposit):
    """Degree 8, two azimuthal orders, 65 frequencies, the route's default basis."""
    from lighthit_amd.compile import compile_axial_slices
    from lighthit_amd.synthetic import plain_small
    field = plain_small().elements.field(0, phase_index=1.339)
    frame = AxisFrame.of(field)
    omega = np.linspace(0.0, 1.2, 65)
    options = dict(azimuthal_degree=2, stencil_threshold_m=0.03, frame=frame,
                   time_deposit=time_deposit)
    expected = oracle.compile_axial_slices(field, 8, omega, **options)
    got = compile_axial_slices(field, 8, omega, **options)
    assert got.summary["slices"] == expected.summary["slices"]
    assert got.summary["time_rows"] == expected.summary["time_rows"]
    _check(got, expected)


def test_spectral_band_and_mask_match_the_band_field(device):
    """A node band from C0, C2, beta on the device equals the oracle on the band's
    photon field (``shower/ge2.py:21-32``); a mask leaves elements out."""
    import torch  # noqa: F401
    from dataclasses import replace
    from lighthit_amd.compile import BasisOptions, DeviceChords, compile_band
    elements = shower(elements=600, seed=5).elements
    rng = np.random.default_rng(2)
    beta = rng.uniform(0.70, 0.9999, len(elements))
    elements = replace(elements, beta=beta, coefficient2=-elements.coefficient0 / beta ** 2)
    mask = rng.uniform(size=len(elements)) < 0.8
    wavelength, phase = 450.0, 1.339
    frame = AxisFrame.of(elements.field(0, phase_index=phase))
    omega = np.linspace(0.0, 1.2, 65)
    chords = DeviceChords(elements)
    got = compile_band(chords, frame, 6, omega, azimuthal_degree=3,
                       options=BasisOptions(stencil_threshold_m=0.03),
                       wavelength_nm=wavelength, phase_index=phase, mask=mask,
                       scale=0.37).to_source()
    taken = mask & (beta * phase > 1)
    part = elements.subset(taken)
    first, second = part.field(0, phase_index=phase), part.field(2, phase_index=phase)
    scale0 = 1.0 / wavelength ** 2
    scale2 = scale0 / phase ** 2
    band = replace(first, photons=scale0 * first.photons - scale2 * second.photons)
    expected = oracle.compile_axial_slices(band, 6, omega, azimuthal_degree=3,
                                           stencil_threshold_m=0.03, frame=frame)
    expected = replace(expected, channels=expected.channels * 0.37)
    _check(got, expected)


def test_detector_scale_shower_compiles_like_the_oracle(device):
    """50,000 chords, the bench event's source."""
    from lighthit_amd.compile import compile_axial_slices
    from lighthit_amd.synthetic import plain_detector_scale
    field = plain_detector_scale().elements.field(0, phase_index=1.339)
    frame = AxisFrame.of(field)
    omega = np.linspace(0.0, 1.2, 65)
    options = dict(azimuthal_degree=2, stencil_threshold_m=0.03, frame=frame)
    expected = oracle.compile_axial_slices(field, 8, omega, **options)
    got = compile_axial_slices(field, 8, omega, **options)
    _check(got, expected)
    # the compile is deterministic: a second run is bit-identical
    again = compile_axial_slices(field, 8, omega, **options)
    assert np.array_equal(again.channels, got.channels)
