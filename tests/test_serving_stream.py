"""The serving stream's events (sizes, placement, reproducibility), its Poisson
arrivals, and host sums that do not depend on the BLAS thread count."""
import numpy as np
from threadpoolctl import threadpool_limits

from lighthit_amd import synthetic
from lighthit_amd.axial import dot3


def test_stream_is_reproducible_and_inside_its_ranges():
    detector = synthetic.detector_scale_array(8)
    a = list(synthetic.serving_stream(detector, 24))
    b = list(synthetic.serving_stream(detector, 24))
    sizes = np.array([len(e) for e in a])
    assert np.array_equal(sizes, [len(e) for e in b])
    assert all(np.array_equal(x.start_m, y.start_m) for x, y in zip(a, b))
    assert sizes.min() >= 500 and sizes.max() <= 200_000
    # log-uniform: the log sizes spread over most of [log 500, log 2e5]
    assert np.ptp(np.log(sizes)) > 0.5 * np.log(400)
    low = detector.positions_m.min(axis=0) - 100 - 10   # a shower is a few m long
    high = detector.positions_m.max(axis=0) + 100 + 10
    for e in a:
        centre = e.start_m.mean(axis=0)
        assert np.all(centre >= low) and np.all(centre <= high)


def test_poisson_arrivals_have_the_rate():
    times = synthetic.poisson_arrivals(200.0, 20_000)
    assert times[0] == 0.0 and np.all(np.diff(times) >= 0)
    assert abs(len(times) / times[-1] - 200.0) < 200.0 * 0.03


def test_dot3_matches_matmul_and_ignores_thread_count():
    rng = np.random.default_rng(3)
    v = rng.normal(size=(100_000, 3))
    axis = rng.normal(size=3)
    np.testing.assert_allclose(dot3(v, axis), v @ axis, rtol=1e-13, atol=1e-13)
    with threadpool_limits(limits=1):
        one = dot3(v, axis)
    assert np.array_equal(one, dot3(v, axis))
