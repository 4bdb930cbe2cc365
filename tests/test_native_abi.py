"""The C-ABI library: it loads, carries the expected ABI version, exports every entry point that
include/mpm_hip.h declares, and the product path fails loudly when the library is missing."""

import ctypes
import os
import re

import pytest

from mpm_plugin_amd import _native as N

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mpm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(mpm_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mpm_create" in names and "mpm_advance" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_abi_version_and_error_string():
    lib = N.lib()
    assert lib.mpm_abi_version() == N.ABI_VERSION
    assert isinstance(lib.mpm_last_error(), bytes)


def test_bindings_cover_header():
    lib = N.lib()
    for n in declared_functions():
        assert getattr(lib, n).restype is not None or n == "mpm_destroy", n


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(N, "_lib", None)
    monkeypatch.setattr(N, "LIB_PATH", str(tmp_path / "libmpm_hip.so"))
    with pytest.raises(N.NativeError, match="missing"):
        N.lib()


def test_null_context_is_rejected():
    lib = N.lib()
    assert lib.mpm_particle_count(None) == 0
    assert lib.mpm_synchronize(None) == N.ERR_VALUE


def test_device_strings():
    from mpm_core.backends.hip.solver import device_ordinal

    assert device_ordinal(None) == 0 and device_ordinal("cuda:0") == 0 and device_ordinal("hip:3") == 3
    with pytest.raises(ValueError):
        device_ordinal("cpu")


def test_struct_layouts_match_the_library():
    """The ctypes structs have the sizes the library's static_asserts fix for the C structs of mpm_hip.h."""
    sizes = {"DomainDesc": 184, "MaterialDesc": 48, "ColliderPose": 132, "AdvanceStats": 24, "SurfaceDesc": 40}
    for name, size in sizes.items():
        assert ctypes.sizeof(getattr(N, name)) == size, name
