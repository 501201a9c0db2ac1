"""The C ABI library loads without a GPU and exports every declared symbol."""
import re
from pathlib import Path

import babelsim_amd
from babelsim_amd import _capi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "bsim.h").read_text()
    return sorted(set(re.findall(r"\b(bsim_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = babelsim_amd.lib()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_symbol_has_a_python_signature():
    bound = set(_capi._DEVICE_SIGNATURES) | set(_capi._VOID) | {
        "bsim_last_error", "bsim_version", "bsim_mesh_generate_hex", "bsim_mesh_from_arrays", "bsim_mesh_prepare",
        "bsim_mesh_sizes", "bsim_mesh_arrays", "bsim_mesh_free", "bsim_server_clock"}
    assert set(declared_symbols()) <= bound


def test_errors_are_reported_not_raised_through_the_abi():
    import pytest
    with pytest.raises(ValueError, match="hex counts"):
        babelsim_amd.MeshData.hex((0, 2, 2))
