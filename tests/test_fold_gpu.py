"""GPU parity of the band-fold re-fit (lh_fold_bundle) against the FP64 oracle."""
import numpy as np
import pytest

from fold_cases import small_kernel
from lighthit_amd.fold import BandFold
from oracle import fold as oracle_fold

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nodes,degree,freq,intervals,rank,subdivide", [
    (4, 4, 9, 16, 4, 1),
    (4, 4, 9, 16, 2, 3),
    (8, 8, 65, 47, 3, 1),      # the detector-scale route: 8 wavelengths folded onto 3 bands
])
def test_fold_matches_oracle(device, nodes, degree, freq, intervals, rank, subdivide):
    from ge2_cases import make_case
    from lighthit_amd.fold import fold_bundle
    bundle, *_ = make_case(7, bands=nodes, degree=degree, freq=freq, intervals=intervals,
                           points=3, receivers=2)
    fold = BandFold.of(small_kernel(nodes=nodes, degree=degree), rank=rank)
    want = oracle_fold.fold_bundle(bundle, fold, subdivide=subdivide)
    got = fold_bundle(bundle, fold, subdivide=subdivide, device=device)
    have = got["coefficients"].cpu().numpy()
    assert have.shape == want["coefficients"].shape
    err = np.abs(have - want["coefficients"]).max() / np.abs(want["coefficients"]).max()
    assert err < 2e-6, err
    np.testing.assert_allclose(got["log_r"], want["log_r"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(got["absorption"], want["absorption"])
    np.testing.assert_allclose(got["speed"], want["speed"])
