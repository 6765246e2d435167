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


def test_tns_reader_fuzz_under_sanitizers(tmp_path):
    """The .tns parser built with AddressSanitizer and UndefinedBehaviorSanitizer
    (host code only, g++) over 150 damaged copies of the two fixtures (bytes
    deleted, inserted, overwritten, files cut short) and a list of edge cases
    (empty file, no final newline, huge or negative indices, inf / nan / 1e999,
    NUL bytes, very long lines): every input either parses or stops with the
    reader's error, and no run reports a memory error or undefined behaviour.
    Buffers the reader hands to its caller are the caller's to free (as with
    the reference's data_buffer), so leak checking is off."""
    import random
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    exe = tmp_path / "tns_dump_san"
    src = os.path.join(ROOT, "diana-tucker_amd", "host", "drivers", "tns_dump.cpp")
    b = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-I", os.path.join(ROOT, "diana-tucker_amd", "host", "include"), "-I",
                        os.path.join(ROOT, "include"), src, "-o", str(exe)], capture_output=True, text=True)
    if b.returncode != 0:
        pytest.skip("sanitizer build unavailable: " + b.stderr[-300:])
    cases = [b"", b"\n", b"# only a comment\n", b"1 2 3 0.5", b"1 2 3 0.5\n1 2 0.5\n", b"0 1 1 1.0\n",
             b"1 2 3 1e999\n", b"1 2 3 nan\n", b"1 2 3 -inf\n", b"1\t2\t3\t4\r\n", b"1 2 3 4 \n",
             b"99999999999999999999999 1 1 1\n", b"4294967297 1 1 1\n", b"1 2 3 4\n\n\n1 2 3 5",
             b"a b c d\n", b"1 2 3 -\n", b"1 2 3 1e\n", b"1 2 3 4\x00\n", b"\x00\x01\x02", b"1 1\n", b"1\n",
             b"1 " * 5000 + b"1\n", b"#\n1 2 3 4\n#x\n2 3 4 5\n"]
    rng = random.Random(5)
    base = [open(os.path.join(ROOT, "tests", "golden", f), "rb").read() for f in ("tiny.tns", "signed_crlf.tns")]
    for _ in range(150):
        d = bytearray(rng.choice(base))
        for _ in range(rng.randint(1, 6)):
            pos = rng.randint(0, max(len(d) - 1, 0))
            op = rng.randint(0, 3)
            if op == 0:
                del d[pos:pos + rng.randint(1, 8)]
            elif op == 1:
                d[pos:pos] = bytes(rng.choice(b"0123456789 .-+eE\t\r\n#x") for _ in range(rng.randint(1, 6)))
            elif op == 2 and d:
                d[pos] = rng.randint(0, 255)
            else:
                d = d[:pos]
        cases.append(bytes(d))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    for k, data in enumerate(cases):
        f = tmp_path / f"case{k}.tns"
        f.write_bytes(data)
        p = subprocess.run([str(exe), str(f)], capture_output=True, env=env, timeout=60)
        err = p.stderr.decode(errors="replace")
        assert "Sanitizer" not in err and "runtime error" not in err, (k, data[:80], err[-2000:])
        assert p.returncode >= 0, (k, data[:80], p.returncode)
        assert p.returncode == 0 or "read_data" in err or "Error" in err, (k, data[:80], p.returncode, err[-500:])
