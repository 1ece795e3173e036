"""CPU pins of the band fold: node choice (product host code) and the re-fit oracle."""
import numpy as np
import pytest

from fold_cases import small_kernel
from lighthit_amd import fold as fold_module
from lighthit_amd.fold import BandFold, KernelPack, frequency_indices, load_folded, save_folded
from oracle import fold as oracle_fold
from oracle import ge2 as oracle_ge2


def test_interpolation_keeps_node_bands_exact():
    """``tests/test_shower_route.py:36-44``."""
    kernel = small_kernel()
    fold = BandFold.of(kernel, rank=2)
    for ell in range(fold.interpolation.shape[0]):
        np.testing.assert_allclose(fold.interpolation[ell][list(fold.bands)], np.eye(2),
                                   atol=1e-9)
    assert fold.residual.max() <= 1.0 + 1e-12
    assert BandFold.of(kernel, rank=4).residual.max() < 1e-8


def test_bands_grow_greedily_beyond_the_exhaustive_limit(monkeypatch):
    """``tests/test_shower_route.py:47-56``."""
    kernel = small_kernel()
    best = BandFold.of(kernel, rank=1)
    monkeypatch.setattr(fold_module, "EXHAUSTIVE_SUBSETS", 1)
    assert BandFold.of(kernel, rank=1).bands == best.bands
    greedy = BandFold.of(kernel, rank=3)
    assert list(greedy.bands) == sorted(set(greedy.bands)) and len(greedy.bands) == 3
    assert set(best.bands) <= set(greedy.bands)
    assert BandFold.of(kernel, rank=4).bands == (0, 1, 2, 3)


def test_route_fold_residual_is_small_at_relativistic_beta():
    """The shower route's three bands of the demonstration quadrature interpolate the
    node-band sources to ~1e-4 for beta > 0.8 (``band_fold.py:25-28``)."""
    kernel = small_kernel(nodes=8, degree=8)
    fold = BandFold.of(kernel, rank=3)
    assert fold.metadata()["worst_residual_design"] < 2e-3
    assert fold.source_residual([0.999], [1.0]) < 1e-3


def test_full_rank_fold_converges_to_the_per_band_apply():
    """``tests/test_shower_route.py:102-138`` through the compile and stage (b)
    oracles: a shower compiled in every band; at full rank only the cubic re-fit
    of the folded radial kernels remains, converging at fourth order."""
    from dataclasses import replace
    from ge2_cases import make_case
    from lighthit_amd.axial import AxisFrame
    from lighthit_amd.sources import SourcePose, SyntheticShower
    from oracle import compile as oracle_compile
    kernel = small_kernel(nodes=4, degree=4)
    fold = BandFold.of(kernel, rank=4)
    bundle, _, positions, looks, alpha, area = make_case(
        5, bands=4, degree=4, azimuthal=2, freq=9, intervals=16, points=3, receivers=5)
    elements = SyntheticShower.gaussian(SourcePose([12., 3., 0.], [0.3, 0.2, 0.93], 0.),
                                        charged_track_length_m=10.0, elements=80,
                                        seed=3).elements
    wavelengths = np.asarray(kernel.wavelength.wavelength_nm, float)
    frame = AxisFrame.of(elements.field(0, phase_index=float(fold.phase_index.max())))
    rng = np.random.default_rng(4)
    direction = rng.normal(size=(5, 3))
    direction /= np.linalg.norm(direction, axis=1)[:, None]
    positions = frame.centre_m + direction * rng.uniform(10.0, 60.0, 5)[:, None]
    sources = []
    for lam in range(4):
        n = float(fold.phase_index[lam])
        first, second = elements.field(0, phase_index=n), elements.field(2, phase_index=n)
        scale0 = 1.0 / wavelengths[lam] ** 2
        field = replace(first, photons=scale0 * first.photons - scale0 / n ** 2 * second.photons)
        sources.append(oracle_compile.compile_axial_slices(
            field, 4, bundle["omega"], azimuthal_degree=2, cell_m=0.5,
            stencil_threshold_m=0.03, frame=frame))
    window = (2.0, 150.0)
    reference = 0
    for lam in range(4):
        band = dict(bundle, coefficients=bundle["coefficients"][lam:lam + 1],
                    absorption=bundle["absorption"][lam:lam + 1],
                    speed=bundle["speed"][lam:lam + 1])
        reference = reference + oracle_ge2.apply_folded(
            band, [sources[lam]], positions, looks, alpha, area * fold.scale[lam], window)
    nodes = [replace(sources[j], channels=sources[j].channels * fold.scale[j]) for j in fold.bands]
    scale = np.abs(reference).max()
    errors = []
    for subdivide in (1, 2, 4):
        folded = oracle_fold.fold_bundle(bundle, fold, subdivide=subdivide)
        folded["coefficients"] = folded["coefficients"].astype(np.complex64)
        result = oracle_ge2.apply_folded(folded, nodes, positions, looks, alpha, area, window)
        errors.append(np.abs(result - reference).max() / scale)
    assert errors[1] < errors[0] / 8 and errors[2] < errors[1] / 8, errors
    assert errors[2] < 2e-4


def test_pack_and_fold_formats_round_trip(tmp_path):
    from ge2_cases import make_case
    bundle, *_ = make_case(2, bands=3, degree=4, freq=9, intervals=6, points=3, receivers=2)
    waves = (411.27, 450.0, 488.73)
    pack = KernelPack()
    pack.add(waves, np.arange(9), bundle)
    pack.save(tmp_path / "pack")
    again = KernelPack.load(tmp_path / "pack")
    sub = again.bundle(waves, 4, [0, 4, 8])
    np.testing.assert_array_equal(sub["coefficients"], bundle["coefficients"][..., [0, 4, 8]])
    np.testing.assert_allclose(sub["omega"], bundle["omega"][[0, 4, 8]])
    with pytest.raises(ValueError):
        again.bundle(waves, 6, [0])
    folded = oracle_fold.fold_bundle(bundle, BandFold.of(small_kernel(nodes=3), rank=2))
    folded["coefficients"] = folded["coefficients"].astype(np.complex64)
    save_folded(folded, tmp_path / "fold")
    loaded = load_folded(tmp_path / "fold")
    np.testing.assert_array_equal(np.asarray(loaded["coefficients"]), folded["coefficients"])
    assert list(frequency_indices(np.linspace(0, 1.2, 65), 65)) == list(range(65))
