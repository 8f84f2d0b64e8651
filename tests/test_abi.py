"""The C ABI without a GPU: the library loads, exports every entry point that
include/numax_hip.h declares, and the ctypes signature table covers exactly
that set.  No compute call is made here."""
import ctypes
import os
import re

import pytest

from numax_amd import _lib

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "numax_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nmx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "nmx_getrf_f64" in names and "nmx_gesv_mpir" in names
    assert len(names) >= 25


def test_signature_table_matches_header():
    assert sorted(_lib.SIGNATURES) == _declared()


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libnumax_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [name for name in _declared() if not hasattr(lib, name)]
    assert not missing, f"not exported: {missing}"


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libnumax_hip.so not built")
def test_status_strings_and_version_need_no_device():
    lib = _lib.load()
    assert lib.nmx_version() > 0
    for code in range(6):
        assert lib.nmx_status_string(code)
