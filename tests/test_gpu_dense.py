"""GPU parity of the dense path: every call goes through libdiana_hip.so's C ABI
and is checked against the FP64 oracle on the same bytes."""
import json
import os

import numpy as np
import pytest

import dtucker as dt
from oracle import dense as D
from oracle import philox

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def gold():
    with open(GOLD) as f:
        g = json.load(f)
    A = np.array(g["tensor_colmajor"]).reshape(g["tensor_shape"], order="F")
    m1 = np.array(g["m1_colmajor"], float).reshape(g["m1_shape"], order="F")
    m2 = np.array(g["m2_colmajor"], float).reshape(g["m2_shape"], order="F")
    return g, A, m1, m2


@pytest.mark.parametrize("dtype", [dt.DT_F64, dt.DT_F32])
def test_ttm1_golden_on_gpu(ctx, dtype):
    g, A, m1, m2 = gold()
    dA = ctx.to_device(A.astype(np.float64 if dtype == dt.DT_F64 else np.float32))
    t = ctx.ttm(dA, ctx.to_device(m1), 1)
    t = ctx.ttm(t, ctx.to_device(m2), 2)
    got = t.download()
    assert np.array_equal(got.ravel(order="F").astype(np.float64), np.array(g["ttm1_colmajor"]))


@pytest.mark.parametrize("dtype", [dt.DT_F64, dt.DT_F32])
def test_gram1_golden_on_gpu(ctx, dtype):
    g, A, _, _ = gold()
    dA = ctx.to_device(A.astype(np.float64 if dtype == dt.DT_F64 else np.float32))
    S = ctx.gram(dA, 1).download()
    assert np.array_equal(S.ravel(order="F"), np.array(g["gram1_colmajor"]))


TTM_CASES = [
    # shape, mode, J -- first, middle, last modes; MFMA (J<=64, aligned) and generic
    ((64, 48, 40), 0, 16), ((64, 48, 40), 1, 16), ((64, 48, 40), 2, 16),
    ((256, 32, 24), 0, 64), ((32, 256, 24), 1, 64), ((32, 24, 256), 2, 64),
    ((96, 80, 16), 0, 8), ((96, 80, 16), 2, 8), ((40, 36, 20), 1, 33), ((40, 36, 20), 0, 70),
    ((16, 256, 64), 1, 16), ((5, 4, 3), 1, 4), ((7, 9, 6, 5), 2, 3), ((128, 64, 8, 8), 3, 48),
    # first mode with K % 128 == 0: for BF16 the persistent kernel with the factor
    # in LDS when B % 32 == 0 (groups of 32 columns: 5 leave waves idle, 1100 give
    # waves two groups), else the direct loads (280, 65); J of one, two and four
    # 16-blocks
    ((128, 40, 7), 0, 33), ((512, 64, 16), 0, 64), ((384, 32, 16), 0, 20), ((256, 64, 8), 0, 7),
    ((256, 13, 5), 0, 7), ((128, 32, 5), 0, 16), ((128, 64, 550), 0, 64),
]


@pytest.mark.parametrize("shape,n,J", TTM_CASES)
def test_ttm_f32_vs_oracle(ctx, shape, n, J):
    print("ttm f32", shape, n, J, flush=True)
    rng = np.random.default_rng(hash((shape, n, J)) % 2**32)
    A = rng.standard_normal(shape).astype(np.float32)
    M = rng.standard_normal((J, shape[n])).astype(np.float32)
    Y = ctx.ttm(ctx.to_device(A), ctx.to_device(M), n).download()
    ref = D.ttm(A.astype(np.float64), M.astype(np.float64), n)
    assert rel(Y, ref) < 2e-6


@pytest.mark.parametrize("shape,n,J", TTM_CASES)
def test_ttm_bf16_vs_oracle(ctx, shape, n, J):
    print("ttm bf16", shape, n, J, flush=True)
    rng = np.random.default_rng(1 + hash((shape, n, J)) % 2**32)
    A = rng.standard_normal(shape).astype(np.float32)
    Ab = dt.bf16_from_f32(A)
    M = rng.standard_normal((J, shape[n])).astype(np.float32)
    dA = ctx.array(shape, dt.DT_BF16).upload(Ab)
    Y = ctx.ttm(dA, ctx.to_device(M), n).download()
    # the kernel rounds M to bf16 too (MFMA path); compare with both operands rounded
    Mb = dt.f32_from_bf16(dt.bf16_from_f32(M)).astype(np.float64)
    ref = D.ttm(dt.f32_from_bf16(Ab).astype(np.float64), Mb, n)
    ref_generic = D.ttm(dt.f32_from_bf16(Ab).astype(np.float64), M.astype(np.float64), n)
    assert min(rel(Y, ref), rel(Y, ref_generic)) < 2e-6


TTMC_CASES = [
    # shape, modes, J -- the first five start with modes {0, 1} in a shape the
    # fused two-mode pass accepts (both FP32 and BF16 for shape[0] % 128 == 0);
    # shape[1] = 64 and 192 leave waves of the workgroup without columns
    ((128, 64, 6), (0, 1), (16, 16)),
    ((128, 128, 3, 4), (1, 0, 3), (5, 30, 2)),
    ((256, 192, 5), (0, 1, 2), (32, 64, 3)),
    ((128, 320, 2), (0, 1), (20, 40)),
    ((64, 64, 2), (0, 1), (8, 40)),
    # not fusable: shape[0] % 64 != 0, modes not {0, 1} first, one product
    ((60, 64, 3), (0, 1), (8, 8)),
    ((128, 64, 7), (2, 0), (3, 4)),
    ((32, 24, 20), (1,), (5,)),
]


def ttmc_oracle(A, Ms, modes):
    Y = A
    for M, n in zip(Ms, modes):
        Y = D.ttm(Y, M, n)
    return Y


@pytest.mark.parametrize("shape,modes,J", TTMC_CASES)
def test_ttmc_f32_vs_oracle(ctx, shape, modes, J):
    """Function::ttmc: chained products; FP32 accumulation over at most two
    contractions of <= 320 terms each gives relative error ~1e-6."""
    print("ttmc f32", shape, modes, J, flush=True)
    rng = np.random.default_rng(11 + len(shape) + sum(J))
    A = rng.standard_normal(shape).astype(np.float32)
    cur = list(shape)
    Ms = []
    for n, j in zip(modes, J):
        Ms.append(rng.standard_normal((j, cur[n])).astype(np.float32))
        cur[n] = j
    Y = ctx.ttmc(ctx.to_device(A), [ctx.to_device(M) for M in Ms], list(modes)).download()
    ref = ttmc_oracle(A.astype(np.float64), [M.astype(np.float64) for M in Ms], modes)
    assert Y.shape == ref.shape
    assert rel(Y, ref) < 3e-6
    # F64 factors give the same FP32 result path
    Y64 = ctx.ttmc(ctx.to_device(A), [ctx.to_device(M.astype(np.float64)) for M in Ms], list(modes))
    assert rel(Y64.download(), ref) < 3e-6


@pytest.mark.parametrize("shape,modes,J", TTMC_CASES)
def test_ttmc_bf16_vs_oracle(ctx, shape, modes, J):
    """BF16 tensor: the product that reads the tensor rounds its factor to
    bf16 (MFMA operand); later products run on the FP32 intermediate."""
    print("ttmc bf16", shape, modes, J, flush=True)
    rng = np.random.default_rng(29 + len(shape) + sum(J))
    A = rng.standard_normal(shape).astype(np.float32)
    Ab = dt.bf16_from_f32(A)
    cur = list(shape)
    Ms = []
    for n, j in zip(modes, J):
        Ms.append(rng.standard_normal((j, cur[n])).astype(np.float32))
        cur[n] = j
    dA = ctx.array(shape, dt.DT_BF16).upload(Ab)
    Y = ctx.ttmc(dA, [ctx.to_device(M) for M in Ms], list(modes)).download()
    Af = dt.f32_from_bf16(Ab).astype(np.float64)
    # fused pair {0, 1}: the mode-0 factor is the bf16 operand, whichever comes first
    first = 0 if set(modes[:2]) == {0, 1} else None
    Mr = []
    for M, n in zip(Ms, modes):
        if (first is not None and n == first) or (first is None and len(Mr) == 0):
            Mr.append(dt.f32_from_bf16(dt.bf16_from_f32(M)).astype(np.float64))
        else:
            Mr.append(M.astype(np.float64))
    if first is not None and modes[0] == 1:
        # mode 0 is contracted first inside the fused pass
        order = [1, 0] + list(range(2, len(modes)))
        ref = ttmc_oracle(Af, [Mr[i] for i in order], [modes[i] for i in order])
    else:
        ref = ttmc_oracle(Af, Mr, modes)
    assert rel(Y, ref) < 3e-6 or not (first is not None and shape[0] % 128 == 0)
    if not (first is not None and shape[0] % 128 == 0):
        # unfused chain: the first product in sequence reads the tensor; its
        # factor is a bf16 MFMA operand on aligned shapes, FP32 on the generic path
        Mr2 = [dt.f32_from_bf16(dt.bf16_from_f32(Ms[0])).astype(np.float64)] + \
              [M.astype(np.float64) for M in Ms[1:]]
        Mr3 = [M.astype(np.float64) for M in Ms]
        assert min(rel(Y, ttmc_oracle(Af, Mr2, modes)), rel(Y, ttmc_oracle(Af, Mr3, modes))) < 3e-6


def test_ttmc_fused_pass_taken(ctx):
    """The {0, 1} pair of a fusable shape runs as one kernel launch."""
    rng = np.random.default_rng(3)
    A = ctx.to_device(rng.standard_normal((128, 64, 4)).astype(np.float32))
    Ms = [ctx.to_device(rng.standard_normal((8, 128)).astype(np.float32)),
          ctx.to_device(rng.standard_normal((8, 64)).astype(np.float32))]
    ctx.timing(True)
    ctx.timing_reset()
    ctx.ttmc(A, Ms, [0, 1])
    names = ctx.timing_list()
    ctx.timing(False)
    assert "ttm2_lead_f32_stream" in names and not any(k.startswith("ttm_m") for k in names)


def test_ttmc_bad_arguments(ctx):
    A = ctx.to_device(np.ones((4, 5, 6), np.float32))
    M = ctx.to_device(np.ones((2, 4), np.float32))
    with pytest.raises(dt.DTError) as e:
        ctx.ttmc(A, [M], [3])
    assert e.value.code == dt.DT_EINVAL


@pytest.mark.parametrize("shape,n", [((64, 48, 40), 0), ((64, 48, 40), 1), ((30, 20, 10), 2),
                                     ((256, 16, 16), 0), ((16, 256, 16), 1), ((7, 5, 130), 2)])
def test_gram_vs_oracle(ctx, shape, n):
    print("gram", shape, n, flush=True)
    rng = np.random.default_rng(7)
    A = rng.standard_normal(shape).astype(np.float32)
    S = ctx.gram(ctx.to_device(A), n).download()
    ref = D.gram(A.astype(np.float64), n)
    assert rel(S, ref) < 2e-6
    S64 = ctx.gram(ctx.to_device(A.astype(np.float64)), n).download()
    assert rel(S64, ref) < 1e-13


@pytest.mark.parametrize("shape,n", [((256, 300, 256), 0), ((256, 300, 256), 1), ((256, 300, 256), 2),
                                     ((520, 160, 420), 0), ((64, 512, 1100), 1), ((128, 512, 1024), 2)])
@pytest.mark.parametrize("bf16", [False, True])
def test_gram_panels_vs_oracle(ctx, shape, n, bf16):
    """Orders >= 256 over >= 65536 columns take the 256-row panel kernel:
    mode 0 (columns contiguous: column layout in LDS), other modes with
    F % 32 == 0 (row layout), orders padded to whole panels (300, 520), one
    and several panels (diagonal and off-diagonal blocks).  FP32 MFMA sums of
    at most a few thousand terms per split, summed in FP64: 1e-5 relative."""
    print("gram panels", shape, n, "bf16", bf16, flush=True)
    rng = np.random.default_rng(17 + n)
    A = rng.standard_normal(shape).astype(np.float32)
    if bf16:
        Ab = dt.bf16_from_f32(A)
        dA = ctx.to_device(Ab)
        A = dt.f32_from_bf16(Ab)
    else:
        dA = ctx.to_device(A)
    S = ctx.gram(dA, n).download()
    ref = D.gram(A.astype(np.float64), n)
    assert rel(S, ref) < 1e-5
    assert np.array_equal(S, S.T)


def test_ttt_except_vs_oracle(ctx):
    rng = np.random.default_rng(8)
    A = rng.standard_normal((6, 7, 5))
    B = rng.standard_normal((6, 3, 5))
    Cm = ctx.ttt_except(ctx.to_device(A), ctx.to_device(B), 1).download()
    assert rel(Cm, D.ttt_except(A, B, 1)) < 1e-13


def test_sumsq(ctx):
    rng = np.random.default_rng(9)
    for n in (1, 1000, 123457):
        x = rng.standard_normal(n)
        assert abs(ctx.sumsq(ctx.to_device(x)) - np.sum(x * x)) < 1e-9 * n
        x32 = x.astype(np.float32)
        assert abs(ctx.sumsq(ctx.to_device(x32)) - np.sum(x32.astype(np.float64) ** 2)) < 1e-9 * n
    # BF16: 8 values per load when n % 8 == 0 (1024, 123456), one per load otherwise
    for n in (7, 1024, 123456, 123457):
        bits = dt.bf16_from_f32(rng.standard_normal(n).astype(np.float32))
        d = ctx.array((n,), dt.DT_BF16).upload(bits)
        xb = dt.f32_from_bf16(bits).astype(np.float64)
        assert abs(ctx.sumsq(d) - np.sum(xb ** 2)) < 1e-9 * n, n


@pytest.mark.parametrize("tA,tB,m,n,k", [(0, 0, 33, 17, 70), (1, 0, 16, 16, 5000),
                                         (0, 1, 64, 40, 8), (1, 1, 9, 70, 31),
                                         (0, 0, 1024, 80, 1024), (1, 0, 80, 80, 1024), (0, 0, 1024, 80, 80),
                                         (0, 1, 130, 65, 200), (1, 1, 129, 67, 3000), (0, 0, 5000, 100, 100)])
def test_gemm_f64(ctx, tA, tB, m, n, k):
    rng = np.random.default_rng(10)
    A = rng.standard_normal((k, m) if tA else (m, k))
    B = rng.standard_normal((n, k) if tB else (k, n))
    Cm = ctx.gemm(tA, tB, ctx.to_device(A), ctx.to_device(B)).download()
    ref = (A.T if tA else A) @ (B.T if tB else B)
    assert rel(Cm, ref) < 1e-14


@pytest.mark.parametrize("m,n", [(64, 8), (1024, 64), (300, 1), (50, 50), (20000, 16), (9001, 29), (4097, 3),
                                 (65537, 32)])
def test_qr_f64(ctx, m, n):
    rng = np.random.default_rng(11)
    A = rng.standard_normal((m, n))
    Q, R = ctx.qr(ctx.to_device(A))
    Q, R = Q.download(), R.download()
    assert np.allclose(Q.T @ Q, np.eye(n), atol=1e-13)
    assert np.allclose(np.triu(R), R)
    assert np.all(np.diag(R) > 0)
    assert rel(Q @ R, A) < 1e-13
    Qn = D.orth_posdiag(A)
    assert np.allclose(Q, Qn, atol=1e-10)


@pytest.mark.parametrize("m,n", [(1024, 64), (700, 40), (300, 17), (5000, 48), (256, 64)])
def test_qr_f64_q_only(ctx, m, n):
    """Q only (no R), 16 < n <= 64: shifted Cholesky-QR3 over many workgroups;
    the same Q as the positive-diagonal QR, also at condition 1e9."""
    rng = np.random.default_rng(m + n)
    for cond in (1.0, 1e9):
        U = np.linalg.qr(rng.standard_normal((m, n)))[0]
        V = np.linalg.qr(rng.standard_normal((n, n)))[0]
        A = U @ np.diag(np.logspace(0, -np.log10(cond), n)) @ V.T
        Q, _ = ctx.qr(ctx.to_device(A), want_r=False)
        Q = Q.download()
        assert np.allclose(Q.T @ Q, np.eye(n), atol=1e-12), cond
        if cond == 1.0:
            assert np.allclose(Q, D.orth_posdiag(A), atol=1e-10)
        else:
            assert D.sin_max_angle(Q, A) < 1e-6


def test_qr_f64_q_only_non_finite_is_enumeric(ctx):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((600, 40))
    A[5, 7] = np.nan
    with pytest.raises(dt.DTError) as e:
        ctx.qr(ctx.to_device(A), want_r=False)
    assert e.value.code == dt.DT_ENUMERIC
    Q, _ = ctx.qr(ctx.to_device(rng.standard_normal((600, 40))), want_r=False)  # the context still works


@pytest.mark.parametrize("n,dup", [(40, (7, 8)), (64, (0, 63)), (20, (3, 3))])
def test_qr_f64_q_only_rank_deficient_is_enumeric(ctx, n, dup):
    """A duplicated column (or, for (3, 3), a zero column) makes the input
    rank-deficient: the Q-only multi-workgroup Cholesky-QR3 reports DT_ENUMERIC
    (the product of its passes' R diagonals falls below 10 n u ||A||_F) instead
    of a basis; an input of condition 1e9 still passes (the shifted-pass case)."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((700, n))
    i, j = dup
    if i == j:
        A[:, i] = 0.0
    else:
        A[:, j] = A[:, i]
    with pytest.raises(dt.DTError) as e:
        ctx.qr(ctx.to_device(A), want_r=False)
    assert e.value.code == dt.DT_ENUMERIC
    U, _ = np.linalg.qr(rng.standard_normal((700, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    B = U @ np.diag(np.logspace(0, -9, n)) @ V.T
    Q, _ = ctx.qr(ctx.to_device(B), want_r=False)
    Qh = Q.download()
    assert np.abs(Qh.T @ Qh - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("m", [500, 6000])
def test_qr_ill_conditioned_uses_shift(ctx, m):
    """Condition 1e9 breaks the plain Cholesky-QR; the shifted pass recovers
    (m = 6000 runs the tall-skinny Gram and row update)."""
    rng = np.random.default_rng(12)
    U = np.linalg.qr(rng.standard_normal((m, 10)))[0]
    V = np.linalg.qr(rng.standard_normal((10, 10)))[0]
    A = U @ np.diag(np.logspace(0, -9, 10)) @ V.T
    Q, R = ctx.qr(ctx.to_device(A))
    Q = Q.download()
    assert np.allclose(Q.T @ Q, np.eye(10), atol=1e-10)
    assert D.sin_max_angle(Q, A) < 1e-6


def test_inverse_f64(ctx):
    rng = np.random.default_rng(13)
    for n in (1, 8, 64, 100):
        A = rng.standard_normal((n, n)) + n * np.eye(n)
        Ci = ctx.inverse(ctx.to_device(A)).download()
        assert np.allclose(Ci @ A, np.eye(n), atol=1e-11)


def test_inverse_singular_reports_enumeric(ctx):
    A = np.ones((4, 4))
    with pytest.raises(dt.DTError) as e:
        ctx.inverse(ctx.to_device(A))
    assert e.value.code == dt.DT_ENUMERIC


@pytest.mark.parametrize("n,r", [(1, 1), (2, 1), (3, 2), (8, 3), (33, 5), (64, 8), (81, 9), (95, 7), (96, 12),
                                 (150, 20), (256, 16), (1024, 64)])
def test_syev_top(ctx, n, r):
    print("syev", n, r, flush=True)
    rng = np.random.default_rng(14)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    lam = np.concatenate([np.logspace(3, 1, r), rng.uniform(0, 1, n - r)])
    S = (Q * lam) @ Q.T
    S = (S + S.T) / 2
    V, w = ctx.syev_top(ctx.to_device(S), r)
    V, w = V.download(), w.download()
    ev, EV = np.linalg.eigh(S)
    assert np.allclose(w, ev[::-1][:r], rtol=1e-10)
    assert D.sin_max_angle(V, EV[:, ::-1][:, :r]) < 1e-8


def test_orth_rows(ctx):
    rng = np.random.default_rng(15)
    A = rng.standard_normal((5, 300))
    out = ctx.orth_rows(ctx.to_device(A)).download()
    assert np.allclose(out @ out.T, np.eye(5), atol=1e-13)
    assert D.sin_max_angle(out.T, A.T) < 1e-12


def test_generators_match_oracle(ctx):
    x = ctx.gen_normal(10001, dt.DT_F64, 20000905, 3).download()
    ref = philox.normal(10001, 20000905, 3)
    assert np.allclose(x, ref, rtol=1e-13, atol=1e-13)
    u = ctx.gen_uniform(777, dt.DT_F64, 5, 9).download()
    assert np.allclose(u, philox.uniform(777, 5, 9), rtol=0, atol=0)


def test_gen_tucker_matches_oracle(ctx):
    shape, ranks = (20, 16, 12), (4, 3, 2)
    g = ctx.gen_tucker(shape, ranks, 1e-2, seed=99, decay=0.9, want_factors=True)
    A = g["A"].download()
    Aref, Uref, _ = D.tucker_model(shape, ranks, 1e-2, 99, decay=0.9)
    assert rel(A, Aref) < 1e-6
    for U, Ur in zip(g["U"], Uref):
        assert np.allclose(U.download(), Ur, atol=1e-12)
    assert abs(g["normA"] - np.linalg.norm(A.astype(np.float64).ravel())) < 1e-6 * g["normA"]


# ---------------------------------------------------------------------------
# HOOI end to end
def run_gpu_hooi(ctx, A, ranks, sweeps, init, U_raw=None, schedule=dt.SCHED_TREE, dtype=dt.DT_F32,
                 flags=0):
    dA = ctx.to_device(A.astype(np.float64 if dtype == dt.DT_F64 else np.float32))
    Uin = [ctx.to_device(u) for u in U_raw] if U_raw is not None else None
    return ctx.hooi_dense(dA, ranks, max_iter=sweeps, init=init, U_init=Uin, schedule=schedule,
                          flags=flags)


@pytest.mark.parametrize("schedule,flags", [
    (dt.SCHED_REFERENCE, 0), (dt.SCHED_TREE, 0),
    (dt.SCHED_REFERENCE, dt.FLAG_FP64_FIT), (dt.SCHED_TREE, dt.FLAG_FP64_FIT),
    (dt.SCHED_REFERENCE, dt.FLAG_FP64_FIT | dt.FLAG_REFERENCE_ALS),
    (dt.SCHED_TREE, dt.FLAG_FP64_FIT | dt.FLAG_REFERENCE_ALS)])
def test_hooi_64cube_reference_random_init(ctx, schedule, flags):
    """64^3, core 8^3, noise 1e-3, ranks 8, 10 sweeps, the reference's own
    random start (srand(20000905), Util::randn, reduced_QR)."""
    shape, ranks = (64, 64, 64), (8, 8, 8)
    g = ctx.gen_tucker(shape, ranks, 1e-3, seed=20000905)
    A = g["A"].download()
    raws, U0 = D.reference_random_factors(shape, ranks, 20000905)
    Gr, Ur, res_ref = D.hooi_als(A.astype(np.float64), ranks, 10, U0)
    out = run_gpu_hooi(ctx, A, ranks, 10, dt.INIT_QR, raws, schedule, flags=flags)
    fit_gpu = 1 - np.array(out["residuals"])
    fit_ref = 1 - np.array(res_ref)
    print("residuals gpu", out["residuals"][-1], "ref", res_ref[-1], flush=True)
    # the true fit of the GPU factors, evaluated by the oracle in FP64
    G_true = D.core(A.astype(np.float64), out["U"])
    res_true = np.sqrt(1 - np.sum(G_true ** 2) / np.sum(A.astype(np.float64) ** 2))
    assert abs((1 - res_true) - fit_ref[-1]) / fit_ref[-1] < 1e-5
    # the residual the GPU reports: FP64-accumulated core -> 1e-5; FP32 core -> 5e-5
    tol = 1e-5 if flags & dt.FLAG_FP64_FIT else 5e-5
    assert np.max(np.abs(fit_gpu - fit_ref) / fit_ref) < tol
    for n in range(3):
        assert D.sin_max_angle(out["U"][n], Ur[n]) < 1e-3
    # core agrees with the oracle core on the GPU factors
    Gg = D.core(A.astype(np.float64), out["U"])
    assert rel(out["G"], Gg) < 1e-5


def test_hooi_64cube_hosvd_and_fp64(ctx):
    shape, ranks = (64, 64, 64), (8, 8, 8)
    g = ctx.gen_tucker(shape, ranks, 1e-3, seed=7)
    A = g["A"].download().astype(np.float64)
    Uh = D.hosvd_factors(A, ranks)
    Gr, Ur, res_ref = D.hooi_als(A, ranks, 10, Uh)
    for dtype in (dt.DT_F32, dt.DT_F64):
        out = run_gpu_hooi(ctx, A, ranks, 10, dt.INIT_HOSVD, dtype=dtype, flags=dt.FLAG_FP64_FIT)
        print("dtype", dtype, "fit gpu", 1 - out["residuals"][-1], "ref", 1 - res_ref[-1], flush=True)
        f_gpu, f_ref = 1 - out["residuals"][-1], 1 - res_ref[-1]
        assert abs(f_gpu - f_ref) / f_ref < (1e-5 if dtype == dt.DT_F32 else 1e-10)
        for n in range(3):
            assert D.sin_max_angle(out["U"][n], Ur[n]) < (1e-3 if dtype == dt.DT_F32 else 1e-8)


def test_hooi_ranks_above_32(ctx):
    """Ranks (40, 48, 64): the factor update runs over many workgroups and its
    order-64 Cholesky is the blocked one (two order-32 wave factorizations);
    fit and factors against the oracle's ALS_ from the same HOSVD start."""
    shape, ranks = (96, 80, 72), (40, 48, 64)
    g = ctx.gen_tucker(shape, ranks, 1e-3, seed=31)
    A = g["A"].download().astype(np.float64)
    Uh = D.hosvd_factors(A, ranks)
    Gr, Ur, res_ref = D.hooi_als(A, ranks, 4, Uh)
    for dtype in (dt.DT_F64, dt.DT_F32):
        out = run_gpu_hooi(ctx, A, ranks, 4, dt.INIT_HOSVD, dtype=dtype, flags=dt.FLAG_FP64_FIT)
        f_gpu, f_ref = 1 - out["residuals"][-1], 1 - res_ref[-1]
        angles = [D.sin_max_angle(out["U"][n], Ur[n]) for n in range(3)]
        print("ranks>32 dtype", dtype, "fit gpu", f_gpu, "ref", f_ref, "angles", angles, flush=True)
        assert abs(f_gpu - f_ref) / f_ref < (1e-10 if dtype == dt.DT_F64 else 1e-5)
        assert max(angles) < (1e-7 if dtype == dt.DT_F64 else 1e-3)


def test_hooi_row_tile_als_matches_oracle(ctx):
    """Modes of 160 to 512 rows at ranks <= 16 take the factor update over
    16-row tiles, one launch per step (200 rows: a ragged last tile; rank 12:
    padded columns); fit and factors against the oracle's ALS_ from the same
    HOSVD start, in FP64 and FP32."""
    shape, ranks = (256, 200, 160), (16, 12, 16)
    g = ctx.gen_tucker(shape, ranks, 1e-3, seed=41)
    A = g["A"].download().astype(np.float64)
    Uh = D.hosvd_factors(A, ranks)
    Gr, Ur, res_ref = D.hooi_als(A, ranks, 4, Uh)
    for dtype in (dt.DT_F64, dt.DT_F32):
        ctx.timing(True)
        ctx.timing_reset()
        out = run_gpu_hooi(ctx, A, ranks, 4, dt.INIT_HOSVD, dtype=dtype, flags=dt.FLAG_FP64_FIT)
        names = ctx.timing_list()
        ctx.timing(False)
        assert "als_tile" in names and "als_fused" not in names
        f_gpu, f_ref = 1 - out["residuals"][-1], 1 - res_ref[-1]
        angles = [D.sin_max_angle(out["U"][n], Ur[n]) for n in range(3)]
        print("row tiles dtype", dtype, "fit gpu", f_gpu, "ref", f_ref, "angles", angles, flush=True)
        assert abs(f_gpu - f_ref) / f_ref < (1e-10 if dtype == dt.DT_F64 else 1e-5)
        assert max(angles) < (1e-7 if dtype == dt.DT_F64 else 1e-3)
        for n in range(3):
            Un = out["U"][n]
            assert np.abs(Un.T @ Un - np.eye(ranks[n])).max() < 1e-12


def test_hooi_ranks_above_32_mode_sizes_multiple_of_128(ctx):
    """Ranks (40, 48, 64) on 256 x 128 x 128: every mode size is a multiple of
    128, so the update's (S Y) R^{-1} steps run on FP64 MFMA (K split over four
    waves) and the start's Q-only QR is the multi-workgroup Cholesky-QR3."""
    shape, ranks = (256, 128, 128), (40, 48, 64)
    g = ctx.gen_tucker(shape, ranks, 1e-3, seed=33)
    A = g["A"].download().astype(np.float64)
    Uh = D.hosvd_factors(A, ranks)
    Gr, Ur, res_ref = D.hooi_als(A, ranks, 3, Uh)
    for dtype in (dt.DT_F64, dt.DT_F32):
        out = run_gpu_hooi(ctx, A, ranks, 3, dt.INIT_HOSVD, dtype=dtype, flags=dt.FLAG_FP64_FIT)
        f_gpu, f_ref = 1 - out["residuals"][-1], 1 - res_ref[-1]
        angles = [D.sin_max_angle(out["U"][n], Ur[n]) for n in range(3)]
        print("ranks>32 x128 dtype", dtype, "fit gpu", f_gpu, "ref", f_ref, "angles", angles, flush=True)
        assert abs(f_gpu - f_ref) / f_ref < (1e-10 if dtype == dt.DT_F64 else 1e-5)
        assert max(angles) < (1e-7 if dtype == dt.DT_F64 else 1e-3)


@pytest.mark.parametrize("ranks,init,flags", [
    ((100, 100, 100), dt.INIT_HOSVD, dt.FLAG_FP64_FIT),
    ((100, 80, 70), dt.INIT_QR, dt.FLAG_FP64_FIT),
    ((100, 100, 100), dt.INIT_QR, dt.FLAG_FP64_FIT | dt.FLAG_REFERENCE_ALS),
    ((96, 100, 65), dt.INIT_HOSVD, 0),
])
def test_hooi_ranks_above_64(ctx, ranks, init, flags):
    """The reference driver's rank (main.cpp:18,24: R = 100) and other ranks
    above 64 on 200 x 160 x 120: the factor update runs its products on the
    tiled FP64 GEMM with the R x R Cholesky (or, with FLAG_REFERENCE_ALS, the
    literal ALS_ inverses) in global memory, the fit correction inverts
    Grams above 64 in global memory, the mode products take the generic
    kernels.  Fit and factors against the oracle's ALS_ from the same start."""
    shape = (200, 160, 120)
    g = ctx.gen_tucker(shape, (100, 100, 100), 1e-2, seed=41)
    A = g["A"].download().astype(np.float64)
    if init == dt.INIT_HOSVD:
        U_raw, U0 = None, D.hosvd_factors(A, ranks)
    else:
        rng = np.random.default_rng(5)
        U_raw = [rng.standard_normal((I, R)) for I, R in zip(shape, ranks)]
        U0 = [D.reduced_qr(u)[0] for u in U_raw]
    _, Ur, res_ref = D.hooi_als(A, ranks, 3, U0)
    for dtype in (dt.DT_F64, dt.DT_F32):
        out = run_gpu_hooi(ctx, A, ranks, 3, init, U_raw=U_raw, dtype=dtype, flags=flags)
        f_gpu, f_ref = 1 - out["residuals"][-1], 1 - res_ref[-1]
        angles = [D.sin_max_angle(out["U"][n], Ur[n]) for n in range(3)]
        print("ranks>64", ranks, "init", init, "flags", flags, "dtype", dtype, "fit gpu", f_gpu, "ref", f_ref,
              "angles", angles, flush=True)
        assert abs(f_gpu - f_ref) / f_ref < (1e-10 if dtype == dt.DT_F64 else 1e-5)
        # ranks below the model's 100 cut inside the signal spectrum, where the
        # eigen-gap is small and three sweeps leave the start's last digits
        # (GPU subspace iteration against numpy's eigh) amplified ~100x
        tol64 = 1e-7 if ranks == (100, 100, 100) else 1e-6
        assert max(angles) < (tol64 if dtype == dt.DT_F64 else 1e-3)
        for n in range(3):
            U = out["U"][n]
            assert np.abs(U.T @ U - np.eye(ranks[n])).max() < 1e-12


def test_hooi_4way_tree_matches_reference_schedule(ctx):
    shape, ranks = (32, 24, 20, 16), (4, 5, 3, 2)
    g = ctx.gen_tucker(shape, ranks, 1e-2, seed=5)
    A = g["A"].download()
    r1 = run_gpu_hooi(ctx, A, ranks, 5, dt.INIT_HOSVD, schedule=dt.SCHED_REFERENCE,
                      flags=dt.FLAG_FP64_FIT)
    r2 = run_gpu_hooi(ctx, A, ranks, 5, dt.INIT_HOSVD, schedule=dt.SCHED_TREE, flags=dt.FLAG_FP64_FIT)
    assert np.allclose(r1["residuals"], r2["residuals"], rtol=1e-6)
    for n in range(4):
        assert D.sin_max_angle(r1["U"][n], r2["U"][n]) < 1e-4
    Uh = D.hosvd_factors(A.astype(np.float64), ranks)
    _, _, res_ref = D.hooi_als(A.astype(np.float64), ranks, 5, Uh)
    assert np.allclose(r2["residuals"], res_ref, rtol=1e-4)


@pytest.mark.parametrize("shape,ranks", [((128, 64, 12, 10), (8, 12, 3, 4)), ((256, 256, 3, 2), (16, 12, 3, 2)),
                                         ((256, 256, 5, 7), (7, 16, 2, 3))])
def test_hooi_4way_tree_fused_pair_matches_reference_schedule(ctx, shape, ranks):
    """Shapes where the tree schedule's second phase contracts modes 0 and 1 in
    one pass (256 x 256 slabs take the whole-column kernel, ragged ranks
    included); the reference schedule runs single products only."""
    g = ctx.gen_tucker(shape, ranks, 1e-2, seed=9)
    A = g["A"].download()
    r1 = run_gpu_hooi(ctx, A, ranks, 4, dt.INIT_HOSVD, schedule=dt.SCHED_REFERENCE,
                      flags=dt.FLAG_FP64_FIT)
    ctx.timing(True)
    ctx.timing_reset()
    r2 = run_gpu_hooi(ctx, A, ranks, 4, dt.INIT_HOSVD, schedule=dt.SCHED_TREE, flags=dt.FLAG_FP64_FIT)
    names = ctx.timing_list()
    ctx.timing(False)
    assert "ttm2_lead_f32_stream" in names
    assert np.allclose(r1["residuals"], r2["residuals"], rtol=1e-6)
    for n in range(4):
        assert D.sin_max_angle(r1["U"][n], r2["U"][n]) < 1e-4
    Uh = D.hosvd_factors(A.astype(np.float64), ranks)
    _, _, res_ref = D.hooi_als(A.astype(np.float64), ranks, 4, Uh)
    assert np.allclose(r2["residuals"], res_ref, rtol=1e-4)


@pytest.mark.parametrize("shape,ranks", [((16, 8, 128, 12), (4, 3, 10, 5)), ((8, 8, 256, 20), (4, 4, 16, 16)),
                                         ((64, 3, 128, 7), (2, 3, 9, 7))])
def test_hooi_4way_tree_trailing_pair_matches_reference_schedule(ctx, shape, ranks):
    """Shapes where the tree schedule's first phase contracts modes 3 and 2 in
    one pass (ranks up to 16, ragged J < 16 included); the reference schedule
    runs single products only."""
    g = ctx.gen_tucker(shape, ranks, 1e-2, seed=13)
    A = g["A"].download()
    r1 = run_gpu_hooi(ctx, A, ranks, 4, dt.INIT_HOSVD, schedule=dt.SCHED_REFERENCE,
                      flags=dt.FLAG_FP64_FIT)
    ctx.timing(True)
    ctx.timing_reset()
    r2 = run_gpu_hooi(ctx, A, ranks, 4, dt.INIT_HOSVD, schedule=dt.SCHED_TREE, flags=dt.FLAG_FP64_FIT)
    names = ctx.timing_list()
    ctx.timing(False)
    assert "ttm2_trail_f32_stream" in names
    assert np.allclose(r1["residuals"], r2["residuals"], rtol=1e-6)
    for n in range(4):
        assert D.sin_max_angle(r1["U"][n], r2["U"][n]) < 1e-4
    Uh = D.hosvd_factors(A.astype(np.float64), ranks)
    _, _, res_ref = D.hooi_als(A.astype(np.float64), ranks, 4, Uh)
    assert np.allclose(r2["residuals"], res_ref, rtol=1e-4)


def test_hooi_bf16_close_to_fp32(ctx):
    shape, ranks = (128, 128, 128), (16, 16, 16)
    g = ctx.gen_tucker(shape, ranks, 1e-3, seed=3, decay=0.93, bf16=True)
    r32 = ctx.hooi_dense(g["A"], ranks, max_iter=4, init=dt.INIT_HOSVD)
    rbf = ctx.hooi_dense(g["A_bf16"], ranks, max_iter=4, init=dt.INIT_HOSVD, A_f32=g["A"],
                         flags=dt.FLAG_FP32_POLISH)
    assert abs((1 - rbf["residual"]) - (1 - r32["residual"])) < 1e-4
    for n in range(3):
        assert D.sin_max_angle(rbf["U"][n], r32["U"][n]) < 0.05


def test_hooi_session_matches_one_shot(ctx):
    shape, ranks = (48, 40, 32), (6, 5, 4)
    g = ctx.gen_tucker(shape, ranks, 1e-2, seed=13)
    one = ctx.hooi_dense(g["A"], ranks, max_iter=3, init=dt.INIT_RANDOM, seed=4)
    s = ctx.hooi_session(g["A"], ranks, init=dt.INIT_RANDOM, seed=4)
    res = [s.sweep() for _ in range(3)]
    out = s.result()
    s.close()
    assert np.allclose(res, one["residuals"], rtol=1e-12)
    assert rel(out["G"], one["G"]) < 1e-12


def test_hooi_session_fp64_fit_without_the_flag(ctx):
    """dt_hooi_fit(fp64 = 1) on a session created without DT_FLAG_FP64_FIT
    (its FP64 chain buffers are made on first use) equals the fit of a session
    created with the flag, and the FP64 oracle's fit of the same factors."""
    shape, ranks = (64, 48, 40), (6, 5, 4)
    g = ctx.gen_tucker(shape, ranks, 1e-3, seed=15, bf16=True)
    fits = []
    for flags in (0, dt.FLAG_FP64_FIT):
        s = ctx.hooi_session(g["A_bf16"], ranks, init=dt.INIT_RANDOM, seed=4, A_f32=g["A"], flags=flags)
        s.sweep()
        fits.append(s.fit(fp64=True))
        U = s.result()["U"]
        s.close()
    assert fits[0] == fits[1]
    A = g["A"].download().astype(np.float64)
    G = A
    for n in range(3):
        G = D.ttm(G, U[n].T, n)
    res = np.sqrt(max(0.0, 1 - np.sum(G * G) / np.sum(A * A)))
    assert abs(fits[0] - res) < 1e-9


def test_hooi_rank_above_size_is_einval(ctx):
    dA = ctx.to_device(np.ones((4, 5, 6), np.float32))
    with pytest.raises(dt.DTError) as e:
        ctx.hooi_dense(dA, (5, 2, 2), max_iter=1, init=dt.INIT_RANDOM)
    assert e.value.code == dt.DT_EINVAL
    # rank 5 of mode 2 above R_0 R_1 = 4: Y_(2) has four columns
    with pytest.raises(dt.DTError) as e:
        ctx.hooi_dense(dA, (2, 2, 5), max_iter=1, init=dt.INIT_RANDOM)
    assert e.value.code == dt.DT_EINVAL


def test_hooi_mode_too_long_for_the_gram_is_enomem(ctx):
    """A 200000 x 2 x 2 tensor: its mode-0 FP64 Gram would need 320 GB, more
    than the device holds, so create fails with DT_ENOMEM naming S (and
    frees what it had allocated: a normal run follows)."""
    dA = ctx.to_device(np.ones((200000, 2, 2), np.float32))
    with pytest.raises(dt.DTError) as e:
        ctx.hooi_dense(dA, (2, 2, 2), max_iter=1, init=dt.INIT_RANDOM)
    print("long mode:", e.value, flush=True)
    assert e.value.code == dt.DT_ENOMEM
    assert "mode Gram S" in str(e.value)
    dA.free()
    g = ctx.gen_tucker((40, 30, 20), (4, 4, 4), 1e-2, seed=3)
    out = ctx.hooi_dense(g["A"], (4, 4, 4), max_iter=2, init=dt.INIT_HOSVD)
    assert out["residual"] < 2e-2


def test_hooi_zero_tensor_is_enumeric(ctx):
    dA = ctx.to_device(np.zeros((8, 8, 8), np.float32))
    with pytest.raises(dt.DTError) as e:
        ctx.hooi_dense(dA, (2, 2, 2), max_iter=1, init=dt.INIT_RANDOM)
    assert e.value.code == dt.DT_ENUMERIC


def test_hooi_256x4_full_size_properties(ctx):
    """256^4 FP32 at its real size (17.2 GB), core 16^4, noise 1e-2,
    HOSVD init, 20 sweeps.  Size-independent checks: the residual reaches the
    noise floor eps*sqrt(1 - prod R/prod I)/sqrt(1+eps^2) and the factors
    recover the true subspaces to the noise level."""
    shape, ranks, eps = (256,) * 4, (16,) * 4, 1e-2
    print("256^4 generate", flush=True)
    g = ctx.gen_tucker(shape, ranks, eps, seed=20000905, want_factors=True)
    print("256^4 hooi", flush=True)
    out = ctx.hooi_dense(g["A"], ranks, max_iter=20, init=dt.INIT_HOSVD)
    pred = eps * np.sqrt(1 - np.prod(ranks) / np.prod(shape)) / np.sqrt(1 + eps * eps)
    print("256^4 residual", out["residual"], "pred", pred, "sweep ms", out["sweeps_ms"] / 20, flush=True)
    assert abs(out["residual"] - pred) / pred < 1e-3
    for n in range(4):
        assert D.sin_max_angle(out["U"][n], g["U"][n].download()) < 5e-3
    del g
