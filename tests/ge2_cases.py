"""Seeded synthetic stage (b) cases shared by the CPU and GPU tests.

The apply is linear in the pack and in the channels, so random packs and
channels of the true shapes exercise every term of equation (*); the
geometry is an axial shower of points and OMs at 2-150 m.
"""
import numpy as np

from lighthit_amd.axial import AxialSource, AxisFrame


def pairs(acceptance_degree):
    lam, mu = [], []
    for a in range(acceptance_degree + 1):
        for m in range(a + 1):
            lam.append(a)
            mu.append(m)
    return np.asarray(lam, np.int64), np.asarray(mu, np.int64)


def make_case(seed, *, bands=3, degree=8, azimuthal=2, acceptance=2, freq=65, points=40,
              receivers=12, intervals=47, r_low=2.0, r_high=150.0, spread_m=60.0):
    rng = np.random.default_rng(seed)
    pl, pm = pairs(acceptance)
    log_r = np.linspace(np.log(r_low), np.log(r_high), intervals + 1)
    # Smooth-ish radial coefficients: O(1) values, powers scaled like a spline.
    coefficients = (rng.normal(size=(bands, intervals, degree + 1, len(pl), 4, freq))
                    + 1j * rng.normal(size=(bands, intervals, degree + 1, len(pl), 4, freq)))
    coefficients[:, :, :, :, 0] *= 3.0
    coefficients = coefficients.astype(np.complex64)
    omega = np.linspace(0., 1.2, freq) if freq > 1 else np.zeros(1)
    bundle = dict(degree=degree, acceptance_degree=acceptance, pair_lambda=pl, pair_mu=pm,
                  omega=omega, phase="flight", coefficients=coefficients, log_r=log_r,
                  absorption=rng.uniform(0.015, 0.03, bands), speed=rng.uniform(0.215, 0.22, bands))
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    helper = np.array([1.0, 0, 0]) if abs(axis[2]) > 0.9 else np.array([0, 0, 1.0])
    first = np.cross(axis, helper)
    first /= np.linalg.norm(first)
    frame = AxisFrame(rng.normal(scale=5.0, size=3), axis, first, np.cross(axis, first))
    z = np.sort(rng.uniform(-3.6, 3.6, points))
    transverse = rng.normal(scale=0.08, size=(points, 2))
    nchan = sum(2 * min(l, azimuthal) + 1 for l in range(degree + 1))
    sources = []
    for band in range(bands):
        channels = (rng.normal(size=(points, nchan, freq))
                    + 1j * rng.normal(size=(points, nchan, freq))) * 50.0
        sources.append(AxialSource(frame, z, transverse, channels, degree, azimuthal,
                                   float(rng.uniform(0.0, 30.0))))
    # OMs around the shower at 10..spread m, some beyond the window on purpose.
    direction = rng.normal(size=(receivers, 3))
    direction /= np.linalg.norm(direction, axis=1)[:, None]
    distance = rng.uniform(8.0, spread_m, receivers)
    positions = frame.centre_m + direction * distance[:, None]
    looks = np.tile([0.0, 0.0, 1.0], (receivers, 1))
    looks[::3] = rng.normal(size=looks[::3].shape)
    looks /= np.linalg.norm(looks, axis=1)[:, None]
    area = np.full(receivers, np.pi * 0.2 ** 2)
    alpha = 2 * np.pi * np.array([0.96666667, 0.3, 0.02666667, 0.0][:acceptance + 1])
    return bundle, sources, positions, looks, alpha, area
