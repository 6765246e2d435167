"""CPU tests pinning the oracle: reference golden vectors, libc's own rand(),
Philox known answers, literal loop transcriptions and closed forms."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import dense as D
from oracle import philox
from oracle.glibc_rand import GlibcRand, util_randn

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")


def load_gold():
    with open(GOLD) as f:
        g = json.load(f)
    A = np.array(g["tensor_colmajor"]).reshape(g["tensor_shape"], order="F")
    m1 = np.array(g["m1_colmajor"], float).reshape(g["m1_shape"], order="F")
    m2 = np.array(g["m2_colmajor"], float).reshape(g["m2_shape"], order="F")
    return g, A, m1, m2


def test_reference_ttm1_golden():
    # tests/testcases/function/distributed/ttm.cpp:7-37
    g, A, m1, m2 = load_gold()
    t = D.ttm(D.ttm(A, m1, 1), m2, 2)
    assert list(t.shape) == g["ttm1_shape"]
    assert np.array_equal(t.ravel(order="F"), np.array(g["ttm1_colmajor"]))


def test_reference_gram1_golden():
    # tests/testcases/function/distributed/gram.cpp:7-19
    g, A, _, _ = load_gold()
    S = D.gram(A, 1)
    assert np.array_equal(S.ravel(order="F"), np.array(g["gram1_colmajor"]))


def test_fixture_matches_per_rank_rule():
    # FunctionDistributedTest.cpp:8-15 with ceil-div blocks: rank 0's block is
    # the 3x2x3 corner, filled 0..17 in column-major order.
    _, A, _, _ = load_gold()
    assert np.array_equal(A[:3, :2, :].ravel(order="F"), np.arange(18.0))


@pytest.mark.parametrize("shape", [(5, 4, 3), (2, 3, 4, 2), (7, 1, 3), (3, 5)])
def test_tenmat_matches_loop_transcription(shape):
    rng = np.random.default_rng(1)
    A = rng.standard_normal(shape)
    flat = list(A.ravel(order="F"))
    for n in range(len(shape)):
        lo = np.array(D.tenmat_loops(flat, list(shape), n))
        assert np.array_equal(D.tenmat(A, n).ravel(order="F"), lo)
        lt = np.array(D.tenmatt_loops(flat, list(shape), n))
        assert np.array_equal(D.tenmatt(A, n).ravel(order="F"), lt)
        # mattten inverts tenmatt
        assert np.array_equal(D.mattten(D.tenmatt(A, n), shape, n), A)


@pytest.mark.parametrize("shape,n,J", [((5, 4, 3), 0, 2), ((5, 4, 3), 1, 6), ((5, 4, 3), 2, 3),
                                       ((3, 4, 2, 5), 2, 4), ((6, 7), 1, 3)])
def test_ttm_brute_force(shape, n, J):
    rng = np.random.default_rng(2)
    A = rng.standard_normal(shape)
    M = rng.standard_normal((J, shape[n]))
    Y = D.ttm(A, M, n)
    # brute force element by element
    out_shape = list(shape)
    out_shape[n] = J
    ref = np.zeros(out_shape)
    for idx in np.ndindex(*out_shape):
        s = 0.0
        for k in range(shape[n]):
            src = list(idx)
            src[n] = k
            s += A[tuple(src)] * M[idx[n], k]
        ref[idx] = s
    assert np.allclose(Y, ref, rtol=1e-13, atol=1e-13)


def test_ttt_except_and_gram_closed_form():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((4, 5, 3))
    B = rng.standard_normal((4, 2, 3))
    C = D.ttt_except(A, B, 1)
    ref = np.einsum("ajc,abc->jb", A, B)
    assert np.allclose(C, ref)
    S = D.gram(A, 2)
    assert np.allclose(S, S.T)
    assert np.all(np.linalg.eigvalsh(S) > -1e-12)


def test_glibc_rand_matches_libc():
    libc = ctypes.CDLL("libc.so.6")
    for seed in (1, 20000905, 0, 424242):
        libc.srand(ctypes.c_uint(seed))
        g = GlibcRand(seed)
        assert [libc.rand() for _ in range(1000)] == [g.rand() for _ in range(1000)]


def test_util_randn_polar_moments():
    g = GlibcRand(7)
    x = np.array([util_randn(g) for _ in range(20000)])
    assert abs(x.mean()) < 0.03 and abs(x.std() - 1) < 0.03


REF_RANDN = os.path.join(os.path.dirname(__file__), "golden", "ref_randn.json")
REF_RANDN_BIN = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "ref_randn")


def _bits(x: float) -> str:
    return "%016x" % int.from_bytes(np.float64(x).tobytes(), "little")


def test_util_randn_matches_reference_fixture():
    # draws of the reference's own Util::randn (src/util.cpp:10-18), compiled
    # from its sources (tests/golden/make_ref_randn.py); bit for bit
    with open(REF_RANDN) as f:
        fx = json.load(f)
    for seed, want in fx["draws"].items():
        g = GlibcRand(int(seed))
        got = [_bits(util_randn(g)) for _ in range(len(want))]
        assert got == want, seed


@pytest.mark.skipif(not os.path.exists(REF_RANDN_BIN), reason="oracle/_ref not built (make -C oracle)")
def test_util_randn_matches_reference_binary_10k():
    import subprocess
    for seed in (20000905, 1, 424242):
        out = subprocess.run([REF_RANDN_BIN, str(seed), "10000"], check=True, capture_output=True,
                             text=True).stdout.split()
        g = GlibcRand(seed)
        assert [_bits(util_randn(g)) for _ in range(10000)] == out, seed


def test_host_util_randn_matches_reference_fixture(tmp_path):
    # the C++ host layer's Util::randn (diana-tucker_amd/host/include/util.hpp)
    # against the same reference draws
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    inc = os.path.join(os.path.dirname(__file__), "..", "diana-tucker_amd", "host", "include")
    src = tmp_path / "r.cpp"
    src.write_text('#include <cstdio>\n#include <cstring>\n#include <cstdint>\n#include "util.hpp"\n'
                   "int main(int c, char **v) { srand((unsigned)strtoul(v[1], 0, 10));\n"
                   "  for (int i = 0; i < 1000; i++) { double x = Util::randn(); uint64_t b;\n"
                   '    memcpy(&b, &x, 8); printf("%016llx\\n", (unsigned long long)b); } }\n')
    exe = tmp_path / "r"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", inc, str(src), "-o", str(exe)], check=True)
    with open(REF_RANDN) as f:
        fx = json.load(f)
    for seed, want in fx["draws"].items():
        out = subprocess.run([str(exe), seed], check=True, capture_output=True, text=True).stdout.split()
        assert out == want[:1000], seed


def test_philox_known_answers():
    # Salmon et al. SC'11 / Random123 known-answer vectors for philox4x32-10
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff, 0xffffffff),
              (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, want in cases:
        got = tuple(int(x) for x in philox.philox4x32_10(*c, *k))
        assert got == want


def test_philox_normal_moments():
    x = philox.normal(200000, 123, 5)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01
    u = philox.uniform(100000, 1, 2)
    assert u.min() > 0 and u.max() <= 1 and abs(u.mean() - 0.5) < 0.01


def test_als_step_is_power_iteration_span():
    """ALS_ returns an orthonormal basis of span(S^k L0), S = Y_(n) Y_(n)^T
    (exact-arithmetic closed form of hooi_als.tpp:7-34)."""
    rng = np.random.default_rng(4)
    Y = rng.standard_normal((9, 4, 5))
    L0 = np.linalg.qr(rng.standard_normal((9, 3)))[0]
    for k in (1, 3, 5):
        Q = D.als_(Y, 0, L0, k)
        S = D.gram(Y, 0)
        ref = np.linalg.matrix_power(S, k) @ L0
        assert np.allclose(Q.T @ Q, np.eye(3), atol=1e-12)
        assert D.sin_max_angle(Q, ref) < 1e-9


def test_sin_max_angle_rotation_invariant():
    rng = np.random.default_rng(5)
    U = np.linalg.qr(rng.standard_normal((20, 4)))[0]
    Rot = np.linalg.qr(rng.standard_normal((4, 4)))[0]
    assert D.sin_max_angle(U, U @ Rot) < 1e-12
    V = np.linalg.qr(rng.standard_normal((20, 4)))[0]
    assert D.sin_max_angle(U, V) > 0.1


def test_hooi_noise_free_exact():
    A, Us, G = D.tucker_model((12, 10, 8), (3, 2, 2), 0.0, seed=11)
    _, U0 = D.reference_random_factors(A.shape, (3, 2, 2), 20000905)
    Gh, Uh, res = D.hooi_als(A.astype(np.float64), (3, 2, 2), 4, U0)
    assert res[-1] < 1e-3   # FP32 storage of X limits this
    for n in range(3):
        assert D.sin_max_angle(Uh[n], Us[n]) < 1e-5


def test_hooi_noise_closed_form():
    """With isotropic noise eps*||X|| the converged residual is close to
    eps*sqrt(1 - prod R / prod I) (noise outside the Tucker subspace)."""
    shape, ranks, eps = (24, 20, 16), (4, 4, 3), 1e-2
    A, Us, G = D.tucker_model(shape, ranks, eps, seed=3)
    Uh = D.hosvd_factors(A.astype(np.float64), ranks)
    _, U, res = D.hooi_als(A.astype(np.float64), ranks, 6, Uh)
    pred = eps * np.sqrt(1 - np.prod(ranks) / np.prod(shape)) / np.sqrt(1 + eps * eps)
    assert abs(res[-1] - pred) / pred < 0.05
    # residual is non-increasing across sweeps (HOOI monotonicity)
    assert all(res[i + 1] <= res[i] + 1e-12 for i in range(len(res) - 1))


def test_reference_random_init_orthonormal_and_repeatable():
    raws, Us = D.reference_random_factors((6, 5), (2, 3), 20000905)
    raws2, _ = D.reference_random_factors((6, 5), (2, 3), 20000905)
    assert all(np.array_equal(a, b) for a, b in zip(raws, raws2))
    for U in Us:
        assert np.allclose(U.T @ U, np.eye(U.shape[1]), atol=1e-12)


# ---- learned start (no reference counterpart; oracle/learned.py) -----------

def test_mlp_oracle_backward_matches_finite_differences():
    """Central differences pin the oracle's hand-written backward (rounding off)."""
    from oracle import learned as Lr
    rng = np.random.default_rng(0)
    din, hid, dout = 6, 5, 3
    n = din * hid + hid + hid * hid + hid + hid * dout + dout
    p = rng.standard_normal(n)
    X = rng.standard_normal((9, din))
    T = rng.standard_normal((9, dout))
    _, g = Lr.loss_grad(p, X, T, din, hid, dout, rounding=False)
    eps = 1e-6
    gn = np.empty(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = eps
        gn[i] = (Lr.loss_grad(p + e, X, T, din, hid, dout, False)[0]
                 - Lr.loss_grad(p - e, X, T, din, hid, dout, False)[0]) / (2 * eps)
    assert np.max(np.abs(g - gn)) < 1e-6 * max(1.0, np.max(np.abs(g)))


def test_bf16_round_matches_binding():
    import sys
    from oracle import learned as Lr
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "diana-tucker_amd", "python"))
    import dtucker as dt
    x = np.random.default_rng(1).standard_normal(1000).astype(np.float32) * 100
    x[:4] = [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 65504.0]  # ties round to even
    assert np.array_equal(Lr.bf16_round(x), dt.f32_from_bf16(dt.bf16_from_f32(x)).astype(np.float64))
    assert Lr.bf16_round(np.float32(1.0 + 2 ** -8)) == 1.0


def test_sketch_features_closed_form():
    """X = sqrt(I)/σ1 · Y V: orthogonal columns with norms sqrt(I) σ_k/σ_1 spanning
    range(Y); the leading R columns recover an exactly rank-R tensor's subspace."""
    from oracle import learned as Lr
    rng = np.random.default_rng(2)
    G = rng.standard_normal((3, 3, 3))
    Us = [np.linalg.qr(rng.standard_normal((I, 3)))[0] for I in (12, 10, 8)]
    A = D.reconstruct(G, Us)
    for n in range(3):
        X = Lr.sketch_features(A, n, 6, 5)
        XtX = X.T @ X
        assert np.allclose(XtX - np.diag(np.diag(XtX)), 0, atol=1e-8 * XtX[0, 0])
        assert np.isclose(XtX[0, 0], A.shape[n])
        assert D.sin_max_angle(np.linalg.qr(X[:, :3])[0], Us[n]) < 1e-8
        # sign convention: largest-magnitude entry of each column positive
        idx = np.argmax(np.abs(X[:, :3]), axis=0)
        assert np.all(X[idx, np.arange(3)] > 0)


# ---- sparse path (oracle/sparse.py) -------------------------------------------

def test_sparse_ttmc_oracle_equals_dense_chain():
    """The per-nonzero Kronecker accumulation equals the dense TTM chain on the
    densified tensor (duplicates summed first: linearity)."""
    from oracle import sparse as S
    rng = np.random.default_rng(3)
    dims = (9, 7, 6)
    idx = [rng.integers(0, d, 80) for d in dims]
    vals = rng.random(80)
    Us = [rng.standard_normal((d, r)) for d, r in zip(dims, (3, 4, 2))]
    A = S.densify(dims, idx, vals)
    for n in range(3):
        Y = S.ttmc(dims, idx, vals, Us, n)
        T = A
        for m in range(3):
            if m != n:
                T = D.ttm(T, Us[m].T, m)
        assert np.allclose(Y, D.tenmat(T, n), rtol=1e-12, atol=1e-12)


def test_sparse_stable_sort_matches_counting_sort():
    """argsort(kind=stable) == the constructor's histogram/scan/scatter in input order."""
    from oracle import sparse as S
    keys = np.random.default_rng(4).integers(0, 17, 500)
    perm, rowptr = S.sort_mode(keys, 17)
    tmp = rowptr[:-1].copy()
    out = np.empty(500, dtype=np.int64)
    for i, k in enumerate(keys):  # sp_tensor.hpp:128-147
        out[tmp[k]] = i
        tmp[k] += 1
    assert np.array_equal(perm, out)


def test_sparse_hooi_oracle_recovers_low_rank():
    """An exactly rank-(2,2,2) tensor stored as COO (all entries): one-step-per-mode
    subspace iteration reaches fit 1 and the true subspaces."""
    from oracle import sparse as S
    rng = np.random.default_rng(5)
    dims, r = (8, 7, 6), 2
    G = rng.standard_normal((r, r, r)) * 5
    Ut = [np.linalg.qr(rng.standard_normal((d, r)))[0] for d in dims]
    A = D.reconstruct(G, Ut)
    idx = [a.ravel(order="F") for a in np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")]
    vals = A.ravel(order="F")
    U0 = [np.linalg.qr(rng.standard_normal((d, r)))[0] for d in dims]
    Us, Gc, res = S.hooi(dims, idx, vals, U0, 25)
    assert res[-1] < 1e-6
    for n in range(3):
        assert D.sin_max_angle(Us[n], Ut[n]) < 1e-6


def test_zipf_cdf_and_merge_small():
    """CDF: monotone, last entry 1, P(0) = 1/H; merged COO: unique keys, sums preserved."""
    from oracle import sparse as S
    c = S.zipf_cdf(1000, 1.1)
    assert np.all(np.diff(c) > 0) and c[-1] == 1.0
    assert np.isclose(c[0], 1.0 / np.sum(np.arange(1, 1001, dtype=float) ** -1.1))
    dims = (50, 40, 30)
    idx, vals = S.zipf_draws(dims, 5000, 1.1, 9)
    uidx, uval = S.merge_duplicates(dims, idx, vals)
    keys = set(zip(*[u.tolist() for u in uidx]))
    assert len(keys) == len(uval) < 5000
    assert np.isclose(float(np.sum(uval, dtype=np.float64)), float(np.sum(vals, dtype=np.float64)), rtol=1e-6)


def test_sparse_ttmc_oracle_equals_dense_chain_nway():
    """4-way and 2-way: the Kronecker rows (first other mode fastest) equal the
    dense TTM chain's unfolding; ttmc_mTR / ttmc_mTL are its products with U_n
    and L (closed forms Y_(n)ᵀ U_n and (Y_(n) L)ᵀ of the dense chain)."""
    from oracle import sparse as S
    rng = np.random.default_rng(13)
    for dims, ranks in (((6, 5, 4, 3), (2, 3, 2, 2)), ((9, 7), (3, 2))):
        nnz = 60
        idx = [rng.integers(0, d, nnz) for d in dims]
        vals = rng.standard_normal(nnz)
        Us = [rng.standard_normal((d, r)) for d, r in zip(dims, ranks)]
        A = S.densify(dims, idx, vals)
        for n in range(len(dims)):
            T = A
            for m in range(len(dims)):
                if m != n:
                    T = D.ttm(T, Us[m].T, m)
            Yd = D.tenmat(T, n)
            assert np.allclose(S.ttmc(dims, idx, vals, Us, n), Yd, rtol=1e-12, atol=1e-12)
            L = S.ttmc_mtr(dims, idx, vals, Us, n)
            assert np.allclose(L, Yd.T @ Us[n], rtol=1e-12, atol=1e-12)
            Lr = rng.standard_normal(L.shape)
            assert np.allclose(S.ttmc_mtl(dims, idx, vals, Us, n, Lr), (Yd @ Lr).T, rtol=1e-12, atol=1e-12)


def test_sparse_merge_first_occurrence_brute_force():
    """Against a dict walk in input order: one entry per coordinate at its first
    occurrence, FP64 sums in input order, survivors in input order; a list
    without duplicates is unchanged."""
    from oracle import sparse as S
    rng = np.random.default_rng(17)
    for dims, nnz in (((5, 4, 3), 200), ((3, 3, 2, 2), 90), ((1000, 900), 50)):
        idx = [rng.integers(0, d, nnz) for d in dims]
        vals = rng.standard_normal(nnz)
        seen = {}
        order = []
        for k in range(nnz):
            c = tuple(int(i[k]) for i in idx)
            if c not in seen:
                seen[c] = 0.0
                order.append(c)
            seen[c] += float(vals[k])
        uidx, uval = S.merge_first_occurrence(idx, vals)
        assert [tuple(int(u[j]) for u in uidx) for j in range(len(uval))] == order
        assert np.array_equal(uval, np.array([seen[c] for c in order]))
    idx = [np.arange(10), np.arange(10)[::-1]]
    uidx, uval = S.merge_first_occurrence(idx, np.arange(10.0))
    assert all(np.array_equal(a, b) for a, b in zip(uidx, idx)) and np.array_equal(uval, np.arange(10.0))


def test_sparse_hooi_oracle_recovers_low_rank_4way():
    """An exactly rank-(2,2,2,2) 4-way tensor stored as COO: fit 1 and the true
    subspaces; the ||L|| the oracle records for the last mode of the last sweep
    equals ||G|| (the core is that mode's L with the updated factor)."""
    from oracle import sparse as S
    rng = np.random.default_rng(19)
    dims, r = (6, 5, 5, 4), 2
    G = rng.standard_normal((r,) * 4) * 5
    Ut = [np.linalg.qr(rng.standard_normal((d, r)))[0] for d in dims]
    A = G
    for n in range(4):
        A = D.ttm(A, Ut[n], n)
    idx = [a.ravel(order="F") for a in np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")]
    vals = A.ravel(order="F")
    U0 = [np.linalg.qr(rng.standard_normal((d, r)))[0] for d in dims]
    ln = []
    Us, Gc, res = S.hooi(dims, idx, vals, U0, 20, lnorms=ln)
    assert res[-1] < 1e-6 and len(ln) == 80
    for n in range(4):
        assert D.sin_max_angle(Us[n], Ut[n]) < 1e-6
    assert np.isclose(np.linalg.norm(Gc), np.linalg.norm(A), rtol=1e-10)


def test_cpu_restatement_matches_oracle(tmp_path):
    """oracle/cpu_hooi.cpp (the C++/OpenMP restatement timed as the CPU
    baseline) gives the numpy oracle's residual per sweep from the same tensor
    and start: 3- and 4-way, ranks below and at a third of the mode sizes."""
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "..", "oracle", "bin", "cpu_hooi")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(os.path.dirname(__file__), "..", "oracle")], check=True)
    rng = np.random.default_rng(3)
    for shape, ranks in (((24, 20, 16), (4, 3, 5)), ((12, 10, 9, 8), (3, 4, 2, 3)), ((40, 35, 30), (6, 6, 6))):
        A = rng.standard_normal(shape)
        U0 = [np.linalg.qr(rng.standard_normal((I, R)))[0] for I, R in zip(shape, ranks)]
        fa, fu = tmp_path / "A.bin", tmp_path / "U.bin"
        A.ravel(order="F").tofile(fa)
        np.concatenate([u.ravel(order="F") for u in U0]).tofile(fu)
        out = subprocess.run([exe, "check", "--shape", ",".join(map(str, shape)), "--ranks", ",".join(map(str, ranks)),
                              "--sweeps", "3", "--tensor", str(fa), "--start", str(fu)],
                             capture_output=True, text=True, check=True).stdout.split()
        _, _, rr = D.hooi_als(A, ranks, 3, U0)
        assert np.allclose([float(x) for x in out], rr, rtol=1e-12), (shape, out, rr)
