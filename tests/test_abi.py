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


def test_tail_args_layout_matches_header(tmp_path):
    """ctypes TailArgs has the offsets and size of struct vox8_tail_args
    (compiled with gcc against include/vox8.h)."""
    import shutil
    import subprocess
    import pytest
    from vox8 import _lib
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    names = [f[0] for f in _lib.TailArgs._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"vox8.h\"\nint main(void) {\n"
                   + "".join(f'  printf("%zu\\n", offsetof(vox8_tail_args, {n}));\n' for n in names)
                   + '  printf("%zu\\n", sizeof(vox8_tail_args));\n  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    want = [getattr(_lib.TailArgs, n).offset for n in names] + [ctypes.sizeof(_lib.TailArgs)]
    assert got == want
