"""The route driven with objects that carry only the reference's attribute sets
(no ``looks``, reference KernelConfig, 321-node default frequency grid forced to
the route's) gives the same response as the package's own kernel."""
import numpy as np
import pytest

from dropin_cases import reference_shaped_demonstration
from route_cases import demonstration_route

pytestmark = pytest.mark.gpu


def test_reference_shaped_kernel_gives_the_same_response(device):
    from lighthit_amd.route import ShowerTransport
    w, kernel, pack, route = demonstration_route(device)
    _, reference_kernel = reference_shaped_demonstration()
    print("drop-in: demonstration event through reference-shaped objects", flush=True)
    dropin = ShowerTransport(reference_kernel, pack, device=device)
    want = route(w.elements)
    got = dropin(w.elements)
    np.testing.assert_array_equal(got.active, want.active)
    np.testing.assert_array_equal(got.charge_components_pe, want.charge_components_pe)
    np.testing.assert_array_equal(got.components_pe, want.components_pe)
    np.testing.assert_array_equal(got.time_origin_ns, want.time_origin_ns)
    assert got.metadata["route"] == "lighthit.shower"


def test_detail_threshold_with_a_receiver_subset(device):
    """transport.py:210-254: with a detail threshold below the threshold, a receiver
    subset is computed at threshold zero and ``active`` then holds the OMs whose
    charge reaches the original threshold."""
    from dataclasses import replace

    from lighthit_amd.route import ShowerTransport
    w, kernel, pack, route = demonstration_route(device)
    ids = np.arange(0, 48, 3)
    # a threshold between the subset's charges
    threshold = float(np.median(route(w.elements).charge_pe[ids]))
    _, reference_kernel = reference_shaped_demonstration()
    reference_kernel.config = replace(reference_kernel.config, threshold_pe=threshold,
                                      detail_threshold_pe=0.01)
    dropin = ShowerTransport(reference_kernel, pack, device=device)
    part = dropin(w.elements, receivers=ids)
    charge = part.charge_pe
    assert np.array_equal(part.active, charge >= threshold)
    assert part.metadata["threshold_pe"] == threshold
    assert part.metadata["detail_threshold_pe"] == 0.01
    assert 0 < part.active.sum() < len(ids)
