"""GPU parity of stage (b), lh_ge2_apply, against the FP64 oracle."""
from dataclasses import replace

import numpy as np
import pytest

from oracle import ge2 as oracle
from ge2_cases import make_case

pytestmark = pytest.mark.gpu

# FP32 contraction against the FP64 oracle: errors relative to the largest
# |value| of the spectrum (random test channels carry both signs).
TOLERANCE = 1e-5


def _check(got, expected, tolerance=TOLERANCE):
    scale = np.abs(expected).max()
    assert scale > 0
    err = np.abs(got - expected).max() / scale
    assert err < tolerance, err


@pytest.mark.parametrize("seed,kw", [
    (0, dict()),                                                     # shower-route shapes
    (1, dict(points=7, receivers=3, freq=9, intervals=12)),          # ragged, small
    (2, dict(degree=16, azimuthal=4, points=30, receivers=5)),       # heavy-event degrees
    (3, dict(acceptance=3, azimuthal=3, degree=6, points=20, receivers=4, bands=2)),
])
def test_full_frequency_apply_matches_oracle(device, seed, kw):
    from lighthit_amd.ge2 import apply_folded
    args = dict(points=165, receivers=34)
    args.update(kw)
    bundle, sources, positions, looks, alpha, area = make_case(seed, **args)
    window = (2.0, 150.0)
    expected = oracle.apply_folded(bundle, sources, positions, looks, alpha, area, window)
    got = apply_folded(bundle, sources, positions, looks, alpha, area, window)
    _check(got, expected)


def test_zero_frequency_apply_matches_oracle(device):
    import torch
    from lighthit_amd import ge2
    bundle, sources, positions, looks, alpha, area = make_case(4, points=165, receivers=48)
    window = (2.0, 150.0)
    zero = dict(bundle, coefficients=np.ascontiguousarray(bundle["coefficients"][..., :1]),
                omega=bundle["omega"][:1])
    zero_sources = [replace(s, channels=np.ascontiguousarray(s.channels[:, :, :1])) for s in sources]
    expected = oracle.apply_folded(zero, zero_sources, positions, looks, alpha, area, window)
    # On the device the full pack and source are resident; one frequency is read by stride.
    pack = ge2.DevicePack(bundle)
    source = ge2.DeviceSource(sources)
    receivers = ge2.DeviceReceivers(source.frame, positions, looks, np.repeat(area[None], 3, 0))
    got = ge2.apply_device(pack, source, receivers, alpha, window, frequencies=1)
    torch.cuda.synchronize()
    _check(got.cpu().numpy(), expected)


def test_different_point_sets_apply_band_by_band(device):
    from lighthit_amd.ge2 import apply_folded
    bundle, sources, positions, looks, alpha, area = make_case(6, points=40, receivers=6, freq=17)
    shifted = [sources[0], replace(sources[1], z_m=sources[1].z_m + 0.01), sources[2]]
    window = (2.0, 150.0)
    expected = oracle.apply_folded(bundle, shifted, positions, looks, alpha, area, window)
    got = apply_folded(bundle, shifted, positions, looks, alpha, area, window)
    _check(got, expected)


def test_window_excludes_points(device):
    from lighthit_amd.ge2 import apply_folded
    bundle, sources, positions, looks, alpha, area = make_case(7, points=50, receivers=8, freq=5)
    window = (20.0, 40.0)
    expected = oracle.apply_folded(bundle, sources, positions, looks, alpha, area, window)
    got = apply_folded(bundle, sources, positions, looks, alpha, area, window)
    assert np.count_nonzero(np.abs(expected).sum(axis=0) == 0) > 0   # some OMs fully outside
    _check(got, expected)


def test_invalid_window_raises(device):
    from lighthit_amd.ge2 import apply_folded
    bundle, sources, positions, looks, alpha, area = make_case(8, points=4, receivers=2, freq=3)
    with pytest.raises(ValueError, match="radial window"):
        apply_folded(bundle, sources, positions, looks, alpha, area, (1.0, 150.0))


SPLIT_CASES = [
    (0, dict()),                                                     # shower-route shapes
    (1, dict(points=7, receivers=3, freq=9, intervals=12)),          # ragged, small
    (2, dict(degree=16, azimuthal=4, points=30, receivers=5)),       # heavy-event degrees
    (3, dict(acceptance=3, azimuthal=3, degree=6, points=20, receivers=4, bands=2)),
    (5, dict(azimuthal=0, degree=4, points=33, receivers=7, freq=3)),
    (9, dict(azimuthal=1, points=50, receivers=40, freq=1)),
]


@pytest.mark.parametrize("seed,kw", SPLIT_CASES)
def test_mfma_split_form_tracks_the_oracle(device, seed, kw):
    """The precision form's multi-frequency apply: bfloat16 hi + lo products on MFMA
    with FP32 accumulation, against the FP64 oracle (every acceptance and
    azimuthal degree built, ragged tiles, one and several frequencies)."""
    from lighthit_amd.ge2 import DevicePack, apply_folded
    args = dict(points=165, receivers=34)
    args.update(kw)
    bundle, sources, positions, looks, alpha, area = make_case(seed, **args)
    window = (2.0, 150.0)
    expected = oracle.apply_folded(bundle, sources, positions, looks, alpha, area, window)
    pack = DevicePack(bundle, coefficient_format="bf16_split")
    got = apply_folded(bundle, sources, positions, looks, alpha, area, window, pack=pack)
    scale = np.abs(expected).max()
    err = float(np.abs(got - expected).max() / scale)
    print(f"mfma split form, case {seed}: max error {err:.2e} of the largest value")
    assert err < 2e-4
    # per OM: the charge-like column (w = 0) within 1e-3 of its own value
    lit = np.abs(expected[0]) > 1e-3 * np.abs(expected[0]).max()
    rel = np.abs(got[0, lit] - expected[0, lit]) / np.abs(expected[0, lit])
    assert rel.max() < 1e-3, rel.max()


def test_mfma_split_form_window_and_point_sets(device):
    from lighthit_amd.ge2 import DevicePack, apply_folded
    bundle, sources, positions, looks, alpha, area = make_case(7, points=50, receivers=8, freq=5)
    pack = DevicePack(bundle, coefficient_format="bf16_split")
    window = (20.0, 40.0)
    expected = oracle.apply_folded(bundle, sources, positions, looks, alpha, area, window)
    got = apply_folded(bundle, sources, positions, looks, alpha, area, window, pack=pack)
    assert np.count_nonzero(np.abs(expected).sum(axis=0) == 0) > 0
    assert np.abs(got - expected).max() < 2e-4 * np.abs(expected).max()
    # OMs with no point inside the window get exactly zero
    empty = np.abs(expected).sum(axis=0) == 0
    assert np.all(got[:, empty] == 0)
    shifted = [sources[0], replace(sources[1], z_m=sources[1].z_m + 0.01), sources[2]]
    window = (2.0, 150.0)
    expected = oracle.apply_folded(bundle, shifted, positions, looks, alpha, area, window)
    got = apply_folded(bundle, shifted, positions, looks, alpha, area, window, pack=pack)
    assert np.abs(got - expected).max() < 2e-4 * np.abs(expected).max()


def test_mfma_split_form_is_repeatable_and_independent_of_grouping(device):
    """Pairs are grouped by radial interval in an order the atomics choose; each
    pair's result and the point-ordered sums do not depend on it."""
    from lighthit_amd.ge2 import DevicePack, apply_folded
    bundle, sources, positions, looks, alpha, area = make_case(11, points=165, receivers=34)
    pack = DevicePack(bundle, coefficient_format="bf16_split")
    window = (2.0, 150.0)
    first = apply_folded(bundle, sources, positions, looks, alpha, area, window, pack=pack)
    for _ in range(3):
        again = apply_folded(bundle, sources, positions, looks, alpha, area, window, pack=pack)
        assert np.array_equal(again, first)
    # an OM alone gives the value it has among the others
    alone = apply_folded(bundle, sources, positions[5:6], looks[5:6], alpha, area[5:6], window,
                         pack=pack)
    assert np.array_equal(alone[:, 0], first[:, 5])
