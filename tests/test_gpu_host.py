"""The C++ host layer (the reference's Tensor/Function/Algorithm API over the
C ABI) on the GPU: the reference's own gtests TTM1/Gram1 rewritten against it,
plus the main.cpp-shaped dense driver."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "diana-tucker_amd", "host", "bin")


def run(exe, *args, timeout=120):
    path = os.path.join(BIN, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    return subprocess.run([path, *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_reference_function_tests_through_host_layer():
    p = run("test_function")
    print(p.stdout, p.stderr)
    assert p.returncode == 0 and "PASSED" in p.stdout


def test_main_driver_prints_reference_lines():
    p = run("main", 60, 50, 40, 8)
    assert p.returncode == 0, p.stderr
    lines = p.stderr.splitlines()
    assert lines[0].startswith("Start Tucker::HOOI_ALS decomposition.. with max_iter = 5")
    res = [float(l.split("=")[-1]) for l in lines if l.startswith("Residual")]
    assert len(res) == 5
    # a Gaussian tensor is far from low rank: the residual stays large and decreases
    assert 0.5 < res[-1] < 1.0 and res[-1] <= res[0] + 1e-9
