"""The heavy single event at full size (10^6 elements, 4608 OMs, 16 wavelength
nodes, source degree 16, azimuthal degree 4, 1560 bins): no oracle finishes
it, so the checks are the properties that hold at any size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def heavy(device):
    from lighthit_amd.fold import KernelPack
    from lighthit_amd.route import (ResidentShower, RouteConfig, RouteKernel, ShowerConfig,
                                    ShowerPrompt, ShowerTransport)
    from lighthit_amd.synthetic import heavy_event, synthetic_pack
    w = heavy_event()
    kernel = RouteKernel(w.medium, w.detector,
                         RouteConfig(relative_time_edges_ns=w.edges_ns, wavelength_nodes=16,
                                     omega_per_ns=w.omega_per_ns,
                                     readout_omega_per_ns=w.readout_omega_per_ns,
                                     scattering_degree=16, source_degree=16,
                                     azimuthal_degree=4))
    pack = KernelPack()
    pack.add(*synthetic_pack(w.medium, kernel.wavelength, degree=16, omega=w.omega_per_ns))
    # the 10 km track lights OMs far out: the selection radius is 1000 m
    route = ShowerTransport(kernel, pack, device=device,
                            config=ShowerConfig(prompt=ShowerPrompt(max_distance_m=1000.0)))
    first = route(ResidentShower(w.elements, device))
    second = route(ResidentShower(w.elements, device))
    from dataclasses import replace
    mixed_route = ShowerTransport(kernel, pack, device=device,
                                  config=replace(route.config, precision="mixed"))
    mixed = mixed_route(ResidentShower(w.elements, device))
    yield w, kernel, first, second, mixed
    # the event's workspaces (tens of GB) go back before the next module
    import torch

    from lighthit_amd import _lib
    del route, mixed_route
    _lib.release_workspaces()
    torch.cuda.empty_cache()


def test_heavy_event_is_repeatable(heavy):
    w, kernel, first, second, _ = heavy
    assert np.array_equal(first.charge_components_pe, second.charge_components_pe)
    assert np.array_equal(first.components_pe, second.components_pe)
    assert np.array_equal(first.active, second.active)


def test_heavy_event_charges_and_mask(heavy):
    w, kernel, r, _, _ = heavy
    n = len(w.detector)
    assert r.charge_components_pe.shape == (n, 3)
    assert r.components_pe.shape == (n, len(w.edges_ns) - 1, 3)
    assert np.isfinite(r.charge_components_pe).all() and np.isfinite(r.components_pe).all()
    assert (r.charge_components_pe[:, :2] >= 0).all()
    assert r.active.sum() > 0
    # active OMs are those at or above threshold
    threshold = kernel.config.threshold_pe
    assert (r.charge_pe[r.active] >= threshold).all()
    # bins are only filled for active OMs, and each order's bins stay within
    # its charge (the window holds part of the light)
    assert not r.components_pe[~r.active].any()
    inside = r.components_pe[r.active].sum(axis=1)
    charge = r.charge_components_pe[r.active]
    assert (inside[:, :2] <= charge[:, :2] * (1 + 1e-6) + 1e-9).all()


def test_heavy_event_precision_form(heavy):
    """The precision form (order >=2 apply as bfloat16 hi + lo and readout as FP16
    hi + lo on MFMA) at the heavy event: every charge, the active set and the
    orders 0 and 1 bins are the FP32 form's; the order >=2 bins differ by the
    split products' rounding only (no FP64 run finishes this event)."""
    w, kernel, r, _, mixed = heavy
    assert np.array_equal(mixed.charge_components_pe, r.charge_components_pe)
    assert np.array_equal(mixed.active, r.active)
    assert np.array_equal(mixed.components_pe[:, :, :2], r.components_pe[:, :, :2])
    a, b = r.components_pe[r.active, :, 2], mixed.components_pe[r.active, :, 2]
    assert np.abs(a - b).sum() <= 1e-4 * np.abs(a).sum()
    s, t = r.spectrum_ge2_pe[1:], mixed.spectrum_ge2_pe[1:]
    assert np.abs(s - t).max() <= 1e-4 * np.abs(s).max()
