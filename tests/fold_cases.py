"""A small kernel-like description (medium, detector, quadrature, degree) for fold tests."""
from types import SimpleNamespace

import numpy as np

from lighthit_amd.model import WavelengthQuadrature
from lighthit_amd.synthetic import baikal_like_detector, synthetic_medium_model


def small_kernel(nodes=4, degree=4):
    return SimpleNamespace(config=SimpleNamespace(source_degree=degree),
                           wavelength=WavelengthQuadrature.gauss_legendre(400.0, 500.0, nodes),
                           detector=baikal_like_detector(clusters=1),
                           medium=synthetic_medium_model())


def smooth_pack(seed, *, bands, degree, freq=9, intervals=16, acceptance=2):
    """Per-band packs of the stage (b) test shapes (flight phase)."""
    from ge2_cases import make_case
    bundle, *_ = make_case(seed, bands=bands, degree=degree, freq=freq, intervals=intervals,
                           acceptance=acceptance, points=5, receivers=3)
    return bundle
