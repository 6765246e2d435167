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


@pytest.mark.parametrize("shape,R", [((60, 50, 40), 8), ((200, 160, 120), 100)])
def test_main_driver_prints_reference_lines(shape, R):
    """main.cpp's workload: srand(20000905), the tensor from Tensor::randn, then
    HOOI_ALS's own random start from the same rand() stream.  The printed
    residuals equal the FP64 oracle's from the identical tensor and start
    (glibc rand() and Util::randn restated); R = 100 is main.cpp:18's rank."""
    import sys
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import dense as D
    from oracle.glibc_rand import GlibcRand, randn_fill
    p = run("main", *shape, R)
    assert p.returncode == 0, p.stderr
    lines = p.stderr.splitlines()
    assert lines[0].startswith("Start Tucker::HOOI_ALS decomposition.. with max_iter = 5")
    res = [float(l.split("=")[-1]) for l in lines if l.startswith("Residual")]
    assert len(res) == 5
    g = GlibcRand(20000905)
    A = np.array(randn_fill(g, int(np.prod(shape)))).reshape(shape, order="F")
    U0 = [D.reduced_qr(np.array(randn_fill(g, I * R)).reshape((I, R), order="F"))[0] for I in shape]
    _, _, rr = D.hooi_als(A, (R,) * 3, 5, U0)
    print("main residuals", res, "oracle", rr, flush=True)
    assert np.allclose(res, rr, atol=2e-6)  # printed with 6 decimals
    check_summary(p.stderr, "Tucker::HOOI_ALS::sweep()", 5)


def test_main_driver_reference_defaults():
    """main.cpp's own workload unchanged (main.cpp:16-24): 500 x 400 x 300
    Gaussian tensor, R = 100, 5 sweeps, FP64.  A Gaussian tensor has no
    low-rank structure, so the residual stays near its start: every printed
    residual is in (0, 1) and no sweep raises it (HOOI's fit is monotone)."""
    p = run("main", timeout=110)
    assert p.returncode == 0, p.stderr[-2000:]
    res = [float(l.split("=")[-1]) for l in p.stderr.splitlines() if l.startswith("Residual")]
    print("main defaults residuals", res, flush=True)
    assert len(res) == 5
    assert all(0.0 < r < 1.0 for r in res)
    assert all(b <= a + 1e-6 for a, b in zip(res, res[1:]))
    check_summary(p.stderr, "Tucker::HOOI_ALS::sweep()", 5)


def check_summary(err, region, calls):
    """The Summary table (src/summary.cpp:98-206 layout): caption row, the
    region with its call count and positive GFlop/s and bandwidth, and the
    per-kernel device times after it."""
    lines = err.splitlines()
    assert any("Time(s)" in l and "GFlop/s" in l and "Bandw.(GB/s)" in l for l in lines)
    k = next(i for i, l in enumerate(lines) if l.startswith("| " + region))
    cells = lines[k + 1].replace("|", " ").split()
    # time, share in nested regions ("Kernel" when none), number, average, GFlop/s, GB/s
    assert int(cells[2]) == calls, cells
    assert float(cells[0]) > 0 and float(cells[4]) > 0 and float(cells[5]) > 0, cells
    assert any(l.startswith("| {Main}") for l in lines)
    assert any(l.startswith("GPU kernels") for l in lines)


def test_spmain_driver_matches_oracle_from_the_reference_start(tmp_path):
    """spmain.cpp's workload through the host layer: read a .tns file, build the
    SpTensor, Sp_HOOI_ALS with truncation 5 from the reference's start (srand(1),
    Tensor {R, I}.randn() per mode); the printed residual equals the sparse
    oracle's from the same start (glibc rand() restated)."""
    import sys
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import dense as D
    from oracle import sparse as S
    from oracle.glibc_rand import GlibcRand, randn_fill
    dims, R, sweeps = (60, 50, 40), 5, 2
    idx, vals = S.gen_zipf(dims, 6000, 1.1, 4)
    vals = np.round(vals.astype(np.float64), 6)  # what the text file holds
    dims = tuple(int(i.max()) + 1 for i in idx)  # read_data: dims = largest index + 1
    path = tmp_path / "small.tns"
    with open(path, "w") as f:
        for k in range(len(vals)):
            f.write(" ".join(str(int(idx[n][k]) + 1) for n in range(3)) + f" {vals[k]:.6f}\n")
    p = run("spmain", path, 1, R, sweeps)
    assert p.returncode == 0, p.stderr
    res = [float(l.split("=")[-1]) for l in p.stderr.splitlines() if l.startswith("Residual")]
    assert len(res) == 1
    g = GlibcRand(1)
    U0 = []
    for I in dims:
        ut = np.array(randn_fill(g, R * I)).reshape((R, I), order="F")  # Tensor {R, I}
        U0.append(D.orth_posdiag(ut.T))
    _, _, rr = S.hooi(dims, idx, vals.astype(np.float32), U0, sweeps)
    print("spmain residual", res[0], "oracle", rr[-1], flush=True)
    assert abs(res[0] - rr[-1]) < 5e-6   # printed with 6 decimals; FP32 TTMc
    check_summary(p.stderr, "SpTucker::Sp_HOOI_ALS::sweeps()", 1)
