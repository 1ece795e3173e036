"""The precision form beside the plain form on the demonstration event
("mixed": the multi-frequency order >=2 applies as bfloat16 hi + lo products
on MFMA, the readout as FP16 hi + lo on MFMA): charge error per order (0, 1,
>=2) as max and median relative error, and the bins' L1 error per order, both
against the FP64 oracle route. The target is a charge within 1e-3 of FP64 on
every OM that carries light of that order, and the same for the bins' L1."""
import json
from dataclasses import replace

import numpy as np
import pytest

from route_cases import demonstration_route

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def forms(device):
    from oracle import route as oracle_route

    from lighthit_amd.route import ShowerTransport
    w, kernel, pack, route = demonstration_route(device)
    plain = route(w.elements)
    forms = {"fp32": plain}
    for form in ("mixed",):
        other = ShowerTransport(kernel, pack, device=device,
                                config=replace(route.config, precision=form))
        forms[form] = other(w.elements)
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, kernel.config.source_degree,
                         np.arange(65))
    want = oracle_route.transport_shower(kernel, w.elements, bundle, None,
                                         config=route.config)
    return forms, want


def error_table(response, want):
    """Per order: max and median relative charge error over the OMs whose FP64
    charge of that order exceeds 1e-6 of the largest, and the bins' L1 error
    over the active OMs relative to their FP64 L1."""
    table = {}
    active = np.flatnonzero(response.active)
    for order, name in ((0, "0"), (1, "1"), (2, ">=2")):
        ref = want["charge"][:, order]
        got = response.charge_components_pe[:, order]
        lit = np.abs(ref) > 1e-6 * np.abs(ref).max()
        rel = np.abs(got[lit] - ref[lit]) / np.abs(ref[lit])
        bins_ref = want["bins"][active, :, order]
        bins_got = response.components_pe[active, :, order]
        l1 = float(np.abs(bins_got - bins_ref).sum() / max(np.abs(bins_ref).sum(), 1e-300))
        table[name] = {"charge_max_rel": float(rel.max()), "charge_median_rel": float(np.median(rel)),
                       "bins_l1_rel": l1, "oms": int(lit.sum())}
    return table


def test_precision_report_meets_the_charge_target(forms):
    responses, want = forms
    report = {form: error_table(r, want) for form, r in responses.items()}
    print("precision report:", json.dumps(report))
    for form in ("fp32", "mixed"):
        for order, row in report[form].items():
            assert row["charge_max_rel"] <= 1e-3, (form, order, row)
            assert row["bins_l1_rel"] <= 1e-3, (form, order, row)


def test_mixed_form_changes_only_the_order_ge2_bins(forms):
    """The precision form keeps every charge (the zero-frequency decisions are
    FP32) and the orders 0 and 1 bins of the plain form; its order >=2 bins and
    spectrum differ by the split products' rounding."""
    responses, _ = forms
    plain, mixed = responses["fp32"], responses["mixed"]
    assert np.array_equal(plain.charge_components_pe, mixed.charge_components_pe)
    assert np.array_equal(plain.active, mixed.active)
    assert np.array_equal(plain.components_pe[:, :, :2], mixed.components_pe[:, :, :2])
    active = np.flatnonzero(plain.active)
    a, b = plain.components_pe[active, :, 2], mixed.components_pe[active, :, 2]
    assert 0 < np.abs(a - b).sum() <= 1e-4 * np.abs(a).sum()
    s, t = plain.spectrum_ge2_pe[1:, active], mixed.spectrum_ge2_pe[1:, active]
    assert 0 < np.abs(s - t).max() <= 1e-4 * np.abs(s).max()
