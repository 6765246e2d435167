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


def run_spmain_vs_oracle(tmp_path, idx, vals, R, sweeps, text_vals=None):
    """Write a .tns file (1-based indices), run spmain on it and compare with the
    sparse oracle from the reference's start (srand(1), Tensor {R, I}.randn()
    per mode, orthonormalised): each sweep's "Calculating iteration" line
    followed by one "||G||_F" line per mode (||L|| of Sp_TTMC_ALS_,
    sp_hooi_als.tpp:51-53), then the core's norm and the residual, all printed
    with six decimals.  Duplicate coordinates are merged before the oracle runs,
    as the SpTensor ingest does; values are FP64 end to end."""
    import sys
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import dense as D
    from oracle import sparse as S
    from oracle.glibc_rand import GlibcRand, randn_fill
    N = len(idx)
    path = tmp_path / "in.tns"
    with open(path, "w") as f:
        for k in range(len(vals)):
            v = text_vals[k] if text_vals is not None else f"{vals[k]:.6f}"
            f.write(" ".join(str(int(idx[n][k]) + 1) for n in range(N)) + f" {v}\n")
    p = run("spmain", path, 1, R, sweeps)
    assert p.returncode == 0, p.stderr
    lines = p.stderr.splitlines()
    dims = tuple(int(i.max()) + 1 for i in idx)  # read_data: dims = largest index + 1
    midx, mvals = S.merge_first_occurrence(idx, vals)
    g = GlibcRand(1)
    U0 = []
    for I in dims:
        ut = np.array(randn_fill(g, R * I)).reshape((R, I), order="F")  # Tensor {R, I}
        U0.append(D.orth_posdiag(ut.T))
    ln = []
    _, Gr, rr = S.hooi(dims, midx, mvals, U0, sweeps, lnorms=ln)
    it = [i for i, l in enumerate(lines) if l.startswith("Calculating iteration")]
    assert len(it) == sweeps
    got = []
    for i in it:
        got += [float(l.split("=")[-1]) for l in lines[i + 1:i + 1 + N]]
        assert all(l.startswith("||G||_F") for l in lines[i + 1:i + 1 + N])
    print("spmain ||L||", got, "oracle", ln, flush=True)
    assert np.allclose(got, ln, rtol=1e-6, atol=2e-6)
    res = [float(l.split("=")[-1]) for l in lines if l.startswith("Residual")]
    assert len(res) == 1
    print("spmain residual", res[0], "oracle", rr[-1], flush=True)
    assert abs(res[0] - rr[-1]) < 2e-6   # printed with 6 decimals; FP64 path
    check_summary(p.stderr, "SpTucker::Sp_HOOI_ALS::sweep()", sweeps)


def test_spmain_driver_matches_oracle_from_the_reference_start(tmp_path):
    """spmain.cpp's workload through the host layer on a 3-way Zipf file."""
    import sys
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import sparse as S
    idx, vals = S.gen_zipf((60, 50, 40), 6000, 1.1, 4)
    vals = np.round(vals.astype(np.float64), 6)  # what the text file holds
    run_spmain_vs_oracle(tmp_path, idx, vals, 5, 2)


def test_spmain_4way_file_signed_values_and_duplicates(tmp_path):
    """A 4-way file whose values carry signs and exponents and whose
    coordinates repeat (merged at ingest), truncation 4, three sweeps."""
    import numpy as np
    rng = np.random.default_rng(8)
    dims, nnz = (30, 25, 20, 12), 5000
    idx = [np.minimum((rng.pareto(1.0, nnz) * d / 6).astype(np.int64), d - 1) for d in dims]
    mant = rng.integers(1, 10 ** 6, nnz)
    expo = rng.integers(-4, 2, nnz)
    sign = np.where(rng.random(nnz) < 0.3, "-", "")
    text = [f"{s}{m}e{e}" for s, m, e in zip(sign, mant, expo)]
    vals = np.array([float(t) for t in text])
    run_spmain_vs_oracle(tmp_path, idx, vals, 4, 3, text_vals=text)
