"""GPU parity of order 1 (lh_order1_*) against the FP64 oracle."""
import numpy as np
import pytest

from lighthit_amd.model import acceptance_alpha
from lighthit_amd.synthetic import plain_small
from oracle import order1 as oracle

pytestmark = pytest.mark.gpu

# FP64 on both sides: agreement to rounding, relative to the largest value
TOLERANCE = 1e-9


def _setup():
    w = plain_small()
    d = w.detector
    tables = oracle.spectral_tables(w.medium, d, w.wavelength)
    el = w.elements
    fastest = float(np.min(tables["group_index"]))
    first = int(np.argmin(el.start_ns))
    origins = el.start_ns[first] + np.linalg.norm(d.positions_m - el.start_m[first],
                                                  axis=1) / (0.299792458 / fastest)
    phase = float(np.sum(tables["s0w"] * tables["phase_index"]) / np.sum(tables["s0w"]))
    return w, d, tables, el, origins, phase


def _close(got, want, tolerance=TOLERANCE):
    scale = np.abs(want).max()
    assert scale > 0
    err = np.abs(np.asarray(got) - np.asarray(want)).max() / scale
    assert err <= tolerance, err


def test_fine_pencils_match_oracle(device):
    from lighthit_amd.compile import DeviceChords
    from lighthit_amd.order1 import DeviceUnitChords, compress
    w, d, tables, el, origins, phase = _setup()
    unit = DeviceUnitChords(DeviceChords(el, device), np.ones(len(el), bool), phase)
    got = compress(unit, 0.5, 96, 192).host()
    want = oracle.compress(el, 1.0 / (el.beta * phase), 0.5, 96, 192)
    assert len(got["w0"]) == len(want["w0"])
    for key in ("position", "direction", "w0", "w2", "time", "major", "var_major",
                "var_minor", "pos_cov"):
        _close(got[key], want[key])


def test_fine_pencils_with_wide_keys_match_oracle(device):
    """5 cm cells make the (cell, pixel) key space exceed 2^32: the node sort
    then runs on 64-bit keys (narrower keys go as uint32)."""
    from lighthit_amd.compile import DeviceChords
    from lighthit_amd.order1 import DeviceUnitChords, compress
    w, d, tables, el, origins, phase = _setup()
    unit = DeviceUnitChords(DeviceChords(el, device), np.ones(len(el), bool), phase)
    grid = unit.grid(0.05, 96, 24)
    assert int(np.prod(list(grid.dims))) * 6 * 96 * 96 >= 2 ** 32
    got = compress(unit, 0.05, 96, 24).host()
    want = oracle.compress(el, 1.0 / (el.beta * phase), 0.05, 96, 24)
    assert len(got["w0"]) == len(want["w0"])
    for key in ("position", "direction", "w0", "w2", "time", "major", "var_major",
                "var_minor", "pos_cov"):
        _close(got[key], want[key])


def test_screen_pencils_match_oracle(device):
    from lighthit_amd.compile import DeviceChords
    from lighthit_amd.order1 import DeviceUnitChords, compress
    w, d, tables, el, origins, phase = _setup()
    unit = DeviceUnitChords(DeviceChords(el, device), np.ones(len(el), bool), phase)
    got = compress(unit, 2.0, 32, 24, parent_face=8).host()
    want = oracle.compress_screen(el, 1.0 / (el.beta * phase), oracle.PromptSettings())
    assert len(got["w0"]) == len(want["w0"])
    assert np.array_equal(got["hierarchy_parent"], want["hierarchy_parent"])
    assert np.array_equal(got["next_node"], want["next_node"])
    for key in ("position", "direction", "w0", "w2", "time", "var_major", "pos_cov"):
        _close(got[key], want[key])


@pytest.mark.parametrize("pool_nodes", [-1, 0], ids=["wide", "single-kernel"])
@pytest.mark.parametrize("screen", [False, True])
def test_grouped_charge_and_bins_match_oracle(device, screen, pool_nodes):
    from lighthit_amd.compile import DeviceChords
    from lighthit_amd.order1 import (DeviceOMs, DeviceOptics, DeviceUnitChords, PromptSettings,
                                     compress, grouped)
    w, d, tables, el, origins, phase = _setup()
    alpha = acceptance_alpha(d)
    settings = PromptSettings()
    unit = DeviceUnitChords(DeviceChords(el, device), np.ones(len(el), bool), phase)
    oms = DeviceOMs(d.positions_m, (-d.orientations), d.effective_area_m2, origins, device)
    optics = DeviceOptics(tables, alpha, w.medium.g, w.edges_ns, settings, device)
    work = np.arange(len(d))
    cone = 1.0 / (el.beta * phase)
    if screen:
        merge = (2e-3, 0.3, 2.0, 0.02)
        pencils = compress(unit, 2.0, 32, 24, parent_face=8)
        want_pencils = oracle.compress_screen(el, cone, oracle.PromptSettings())
    else:
        merge = (2e-3, 0.08, 0.5, 0.02)
        pencils = compress(unit, 0.5, 96, 192)
        want_pencils = oracle.compress(el, cone, 0.5, 96, 192)
    charge, bins = grouped(pencils, oms, work, optics, merge, compute_bins=not screen,
                           pool_nodes=pool_nodes)
    want_charge, want_bins, _, _ = oracle.grouped(
        d.positions_m, (-d.orientations), d.effective_area_m2, origins, work, want_pencils, tables, alpha,
        w.medium.g, w.edges_ns, oracle.PromptSettings(), merge, compute_bins=not screen)
    _close(charge.cpu().numpy(), want_charge)
    if not screen:
        _close(bins.cpu().numpy(), want_bins)


def test_wide_and_single_kernel_batches_agree(device):
    """The two forms of the batch kernels sum the same terms in the same order."""
    from lighthit_amd.compile import DeviceChords
    from lighthit_amd.order1 import (DeviceOMs, DeviceOptics, DeviceUnitChords, PromptSettings,
                                     compress, grouped)
    w, d, tables, el, origins, phase = _setup()
    alpha = acceptance_alpha(d)
    unit = DeviceUnitChords(DeviceChords(el, device), np.ones(len(el), bool), phase)
    oms = DeviceOMs(d.positions_m, (-d.orientations), d.effective_area_m2, origins, device)
    optics = DeviceOptics(tables, alpha, w.medium.g, w.edges_ns, PromptSettings(), device)
    pencils = compress(unit, 0.5, 96, 192)
    merge = (2e-3, 0.08, 0.5, 0.02)
    work = np.arange(len(d))
    wide = [t.cpu().numpy() for t in grouped(pencils, oms, work, optics, merge)]
    single = [t.cpu().numpy() for t in grouped(pencils, oms, work, optics, merge, pool_nodes=0)]
    for a, b in zip(wide, single):
        _close(a, b, 1e-13)


def test_grouped_in_parts_matches_one_call(device, monkeypatch):
    """A long OM list runs in parts of bounded size; every OM gets the charge and
    bins of the single call (segment sums may round differently)."""
    from lighthit_amd import order1
    from lighthit_amd.compile import DeviceChords
    from lighthit_amd.order1 import (DeviceOMs, DeviceOptics, DeviceUnitChords, PromptSettings,
                                     compress, grouped)
    w, d, tables, el, origins, phase = _setup()
    alpha = acceptance_alpha(d)
    unit = DeviceUnitChords(DeviceChords(el, device), np.ones(len(el), bool), phase)
    oms = DeviceOMs(d.positions_m, (-d.orientations), d.effective_area_m2, origins, device)
    optics = DeviceOptics(tables, alpha, w.medium.g, w.edges_ns, PromptSettings(), device)
    pencils = compress(unit, 0.5, 96, 192)
    merge = (2e-3, 0.08, 0.5, 0.02)
    work = np.arange(len(d))
    whole = [t.cpu().numpy() for t in grouped(pencils, oms, work, optics, merge)]
    monkeypatch.setattr(order1, "MAX_GROUPED_ENTRIES", 7 * int(pencils.count))
    parts = [t.cpu().numpy() for t in grouped(pencils, oms, work, optics, merge)]
    for a, b in zip(whole, parts):
        _close(b, a, 1e-12)


def test_spectral_order1_with_global_screen_matches_oracle(device):
    from lighthit_amd.compile import DeviceChords
    from lighthit_amd.order1 import DeviceOMs, spectral_order1
    w, d, tables, el, origins, phase = _setup()
    alpha = acceptance_alpha(d)
    rng = np.random.default_rng(3)
    q0 = rng.uniform(0.0, 0.02, len(d))
    extra = rng.uniform(0.0, 0.005, len(d))
    oms = DeviceOMs(d.positions_m, (-d.orientations), d.effective_area_m2, origins, device)
    work = np.arange(len(d))
    got = spectral_order1(DeviceChords(el, device), oms, work, tables, alpha, w.medium.g,
                          w.edges_ns, 0.01, q0, extra)
    want = oracle.spectral_order1(el, d.positions_m, (-d.orientations), d.effective_area_m2, origins, work,
                                  tables, alpha, w.medium.g, w.edges_ns, 0.01, q0, extra)
    assert np.array_equal(got[2], want[2])
    _close(got[0], want[0])
    _close(got[1], want[1])
    assert abs(got[0].sum() - 50.68) < 0.01


def test_grouped_keys_wider_than_32_bits_match_oracle(device, monkeypatch):
    """A 1e-5 ns delay step makes the per-OM key space exceed 2^32: the grouped
    pass then stores and sorts 64-bit keys (narrower keys go as uint32)."""
    from lighthit_amd import _lib
    from lighthit_amd.compile import DeviceChords
    from lighthit_amd.order1 import (DeviceOMs, DeviceOptics, DeviceUnitChords, PromptSettings,
                                     compress, grouped)
    w, d, tables, el, origins, phase = _setup()
    alpha = acceptance_alpha(d)
    unit = DeviceUnitChords(DeviceChords(el, device), np.ones(len(el), bool), phase)
    oms = DeviceOMs(d.positions_m, (-d.orientations), d.effective_area_m2, origins, device)
    optics = DeviceOptics(tables, alpha, w.medium.g, w.edges_ns, PromptSettings(), device)
    pencils = compress(unit, 0.5, 96, 192)
    merge = (2e-3, 0.08, 1e-5, 0.02)
    work = np.arange(len(d))
    seen = []
    call = _lib.call

    def spy(name, *args):
        if name == "lh_order1_group_charge":
            seen.append(int(args[5]))      # the key bits of the compact per-OM keys
        return call(name, *args)
    monkeypatch.setattr(_lib, "call", spy)
    charge, bins = grouped(pencils, oms, work, optics, merge)
    assert seen and seen[0] + int(len(work) - 1).bit_length() > 32
    want_pencils = oracle.compress(el, 1.0 / (el.beta * phase), 0.5, 96, 192)
    want_charge, want_bins, _, _ = oracle.grouped(
        d.positions_m, (-d.orientations), d.effective_area_m2, origins, work, want_pencils, tables, alpha,
        w.medium.g, w.edges_ns, oracle.PromptSettings(), merge, compute_bins=True)
    _close(charge.cpu().numpy(), want_charge)
    _close(bins.cpu().numpy(), want_bins)
