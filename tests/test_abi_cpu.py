"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares (no compute calls are made here)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "diana-tucker_amd", "libdiana_hip.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|const char \*)\s*\*?\s*(dt_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "dt_ttm" in names and "dt_hooi_dense" in names
    assert len(names) >= 30


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_binding_covers_declared_symbols():
    import dtucker
    names = set(declared_functions())
    assert set(dtucker.exported_symbols()) <= names


def test_tns_reader_fixture():
    """read_data rules (src/read_data.cpp:13-33, 80-127): space-separated,
    1-based indices, ndim = spaces on the first line, dims = max index + 1."""
    import subprocess
    exe = os.path.join(ROOT, "diana-tucker_amd", "host", "bin", "tns_dump")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "diana-tucker_amd", "host")], check=True)
    out = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "tiny.tns")], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    assert out[0] == "3 4"
    assert out[1] == "dims 4 3 3"
    assert out[2:6] == ["0 0 0 1.5", "1 2 0 0.25", "3 0 1 2", "2 2 2 10.125"]
