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


def tns_dump():
    import subprocess
    exe = os.path.join(ROOT, "diana-tucker_amd", "host", "bin", "tns_dump")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "diana-tucker_amd", "host")], check=True)
    return exe


def test_tns_reader_fixture():
    """read_data rules (src/read_data.cpp:80-127): 1-based indices, ndim = fields
    on the first line minus one, dims = max index + 1."""
    import subprocess
    out = subprocess.run([tns_dump(), os.path.join(ROOT, "tests", "golden", "tiny.tns")], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    assert out[0] == "3 4"
    assert out[1] == "dims 4 3 3"
    assert out[2:6] == ["0 0 0 1.5", "1 2 0 0.25", "3 0 1 2", "2 2 2 10.125"]


def test_tns_reader_signs_exponents_crlf_tabs():
    """Values with signs and exponents, CRLF line ends, a tab separator, a
    comment and a blank line (the reference's scanner, read_data.cpp:13-33,
    accepts only digits and '.', so '-1.5' would shift every later field)."""
    import subprocess
    out = subprocess.run([tns_dump(), os.path.join(ROOT, "tests", "golden", "signed_crlf.tns")], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    assert out[0] == "3 5"
    assert out[1] == "dims 4 3 3"
    assert out[2:7] == ["0 0 0 -1.5", "1 2 0 0.0025000000000000001", "3 0 1 -700", "2 2 2 0.125", "0 0 0 0.5"]


@pytest.mark.parametrize("text,msg", [("1 1 1 0.5\n0 2 1 1.0\n", "index 0"),
                                      ("1 1 1 0.5\n2 2 1.0\n", "3 fields, expected 4"),
                                      ("1 1 1 0.5\n2 x 1 1.0\n", "not an unsigned integer"),
                                      ("1 1 1 0.5\n2 2 1 1.0.0\n", "not a number"),
                                      ("1 1 1 0.5\n2 -2 1 1.0\n", "not an unsigned integer")])
def test_tns_reader_rejects_malformed_lines(tmp_path, text, msg):
    """A malformed line stops the reader with an error naming the file and line
    (status -3, the reference's error() convention)."""
    import subprocess
    f = tmp_path / "bad.tns"
    f.write_text(text)
    p = subprocess.run([tns_dump(), str(f)], capture_output=True, text=True)
    assert p.returncode != 0
    assert msg in p.stderr and "bad.tns:2" in p.stderr
