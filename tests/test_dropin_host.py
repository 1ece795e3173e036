"""The route accepts the reference's own kernel and data objects (CPU side)."""
import json

import numpy as np
import pytest

from dropin_cases import reference_shaped_demonstration


def test_route_forces_its_frequency_grids_on_a_reference_kernel():
    from lighthit_amd.fold import KernelPack
    from lighthit_amd.route import FREQUENCIES_PER_NS, READOUT_PER_NS, ShowerTransport
    w, kernel = reference_shaped_demonstration()
    route = ShowerTransport(kernel, KernelPack(), device="cpu")
    config = route.kernel.config
    np.testing.assert_array_equal(config.omega_per_ns, FREQUENCIES_PER_NS)
    np.testing.assert_array_equal(config.readout_omega_per_ns, READOUT_PER_NS)
    np.testing.assert_array_equal(config.relative_time_edges_ns, w.edges_ns)
    assert route.kernel.detector is kernel.detector
    assert route.kernel.wavelength is kernel.wavelength


def test_shower_data_manifest_and_search(tmp_path, monkeypatch):
    from lighthit_amd.route import ShowerData
    cache = tmp_path / "demo-cache"
    (cache / "ge2-pack").mkdir(parents=True)
    (cache / "shower-data.json").write_text(json.dumps({
        "rte_cache": ".", "ge2_pack": "ge2-pack", "kernel_settings": {"threshold_pe": 0.02}}))
    monkeypatch.delenv("LIGHTHIT_SHOWER_DATA", raising=False)
    found = ShowerData.find(cache)
    assert found.ge2_pack == (cache / "ge2-pack").resolve()
    assert found.rte_cache == cache.resolve()
    assert found.kernel_settings == {"threshold_pe": 0.02}
    monkeypatch.setenv("LIGHTHIT_SHOWER_DATA", str(cache))
    assert ShowerData.find(None).manifest == (cache / "shower-data.json").resolve()
    monkeypatch.delenv("LIGHTHIT_SHOWER_DATA")
    with pytest.raises(FileNotFoundError, match="searched"):
        ShowerData.find(tmp_path / "elsewhere")
    with pytest.raises(FileNotFoundError, match="incomplete"):
        ShowerData.of(cache, cache / "missing-pack")


def test_data_kernel_settings_reach_the_route_kernel(tmp_path):
    from lighthit_amd.route import ShowerData, ShowerTransport
    w, kernel = reference_shaped_demonstration()
    (tmp_path / "pack").mkdir()
    data = ShowerData(None, tmp_path / "pack", kernel_settings={"threshold_pe": 0.05})
    route = ShowerTransport(kernel, data, device="cpu")
    assert route.kernel.config.threshold_pe == 0.05
    assert kernel.config.threshold_pe == 0.01


def test_looks_are_minus_orientations_of_a_reference_detector():
    from lighthit_amd.model import look_directions
    w, kernel = reference_shaped_demonstration()
    looks = look_directions(kernel.detector)
    np.testing.assert_array_equal(looks, -kernel.detector.orientations)
    assert not hasattr(kernel.detector, "looks")
