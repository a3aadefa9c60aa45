"""The C-ABI library loads and exports every symbol declared in include/vox8.h
(no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

from conftest import ROOT


def test_library_exports_header_symbols():
    from vox8 import _lib
    header = (ROOT / "include" / "vox8.h").read_text()
    declared = set(re.findall(r"\b(vox8_[a-z0-9_]+)\s*\(", header))
    assert len(declared) >= 20
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) | {"vox8_last_error"} == declared
    assert lib.vox8_version() == 100
