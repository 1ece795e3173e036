"""The C-ABI library loads (without a GPU) and exports every declared entry point."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "lighthit_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "lh_readout_bins" in names and "lh_ge2_apply" in names


def test_library_exports_every_declared_symbol():
    from lighthit_amd import _lib
    if not _lib.LIBRARY.exists():
        pytest.fail("liblighthit_amd.so is not built (run python -m lighthit_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIBRARY))
    missing = [name for name in declared() if not hasattr(lib, name)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    from lighthit_amd import _lib
    import lighthit_amd.ge2  # noqa: F401  registers its entry points
    import lighthit_amd.ballistic  # noqa: F401
    import lighthit_amd.compile  # noqa: F401
    import lighthit_amd.order1  # noqa: F401
    import lighthit_amd.fold  # noqa: F401
    import lighthit_amd.route  # noqa: F401
    import lighthit_amd.near_field  # noqa: F401
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_abi_version_and_error_state():
    from lighthit_amd import _lib
    lib = _lib.load()
    assert lib.lh_abi_version() == 1
    lib.lh_clear_error()
    assert lib.lh_last_error() == b""


def test_argument_errors_map_to_value_error():
    # A validation failure returns before any device work: safe without a GPU.
    from lighthit_amd import _lib
    with pytest.raises(ValueError, match="n_bins > 0"):
        _lib.call("lh_readout_bins", None, None, None, None, 0, 65, 1, None, None)
