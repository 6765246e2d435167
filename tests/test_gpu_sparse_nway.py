"""Sparse tensors of any order and rank on the GPU: the general TTMc
(ttmc_mTR / ttmc_mTL at the C ABI), N-way sparse HOOI in FP32 and FP64,
the per-sweep ||L|| of the session API, and the ingest merge of duplicate
coordinates -- all against oracle/sparse.py."""
import numpy as np
import pytest

import dtucker as dt
from oracle import dense as D
from oracle import sparse as S

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def coo(rng, dims, nnz, hot=None):
    """Skewed COO, duplicates merged, optionally one mode-0 slice with `hot`
    nonzeros (longer than a chunk of the general TTMc)."""
    idx = [np.minimum((rng.pareto(1.0, nnz) * d / 8).astype(np.int64), d - 1) for d in dims]
    if hot:
        idx[0][:hot] = 1
    vals = rng.random(nnz)
    uidx, uval = S.merge_first_occurrence(idx, vals)
    return [u.astype(np.int32) for u in uidx], uval


CASES = [((40, 30, 20, 10), (3, 4, 5, 2), 6000, None), ((300, 200, 150), (24, 32, 20), 20000, 3000),
         ((500, 400), (9, 9), 5000, None), ((12, 10, 8, 6, 5), (2, 3, 2, 3, 2), 4000, None),
         ((200, 150, 100), (16, 16, 16), 30000, 9000)]


@pytest.mark.parametrize("dims,ranks,nnz,hot", CASES)
@pytest.mark.parametrize("f64", [False, True])
def test_ttmc_mtr_mtl_vs_oracle(ctx, dims, ranks, nnz, hot, f64):
    """L = Y_(n)ᵀ U_n and (Y_(n) L)ᵀ for every mode: FP64 values to 1e-12; FP32
    values to 1e-5 (Kronecker rows summed in FP32 over up to 3000 nonzeros of a
    slice: ~sqrt(3000) * 6e-8 = 3e-6 relative; the products with U_n or L in FP64)."""
    print("mtr/mtl", dims, ranks, "f64", f64, flush=True)
    rng = np.random.default_rng(len(dims) + sum(ranks))
    idx, vals = coo(rng, dims, nnz, hot)
    Us = [np.linalg.qr(rng.standard_normal((d, r)))[0] for d, r in zip(dims, ranks)]
    v = vals if f64 else vals.astype(np.float32)
    sp = dt.SpTensor.from_host(ctx, dims, idx, v, f64=f64)
    tol = 1e-12 if f64 else 1e-5
    try:
        for n in range(len(dims)):
            L = sp.ttmc_mtr(n, Us)
            Lr = S.ttmc_mtr(dims, idx, v.astype(np.float64), Us, n)
            assert rel(L, Lr) < tol, (n, rel(L, Lr))
            Rt = sp.ttmc_mtl(n, Us, Lr)
            Rr = S.ttmc_mtl(dims, idx, v.astype(np.float64), Us, n, Lr)
            assert Rt.shape == (ranks[n], dims[n])
            assert rel(Rt, Rr) < tol, (n, rel(Rt, Rr))
    finally:
        sp.close()


@pytest.mark.parametrize("dims,ranks,nnz,hot", CASES)
@pytest.mark.parametrize("f64", [False, True])
def test_sparse_hooi_any_order_vs_oracle(ctx, dims, ranks, nnz, hot, f64):
    """Three sweeps from the same start through the session API: residuals,
    every mode's ||L|| per sweep (what Sp_TTMC_ALS_ prints), the core's norm
    and the factors by principal angles."""
    print("sparse hooi", dims, ranks, "f64", f64, flush=True)
    rng = np.random.default_rng(7 + len(dims))
    idx, vals = coo(rng, dims, nnz, hot)
    v = vals if f64 else vals.astype(np.float32)
    U0 = [D.orth_posdiag(rng.standard_normal((d, r))) for d, r in zip(dims, ranks)]
    sp = dt.SpTensor.from_host(ctx, dims, idx, v, f64=f64)
    try:
        s = sp.hooi_session(ranks, U_init=U0)
        res, ln = [], []
        for _ in range(3):
            r, l = s.sweep()
            res.append(r)
            ln.extend(l)
        out = s.result()
        s.close()
    finally:
        sp.close()
    lr = []
    Ur, Gr, rr = S.hooi(dims, idx, v.astype(np.float64), U0, 3, lnorms=lr)
    print("residuals", res, rr, flush=True)
    rt, at = (1e-10, 1e-7) if f64 else (1e-5, 1e-3)
    assert np.allclose(res, rr, rtol=rt)
    assert np.allclose(ln, lr, rtol=rt)
    assert np.isclose(np.linalg.norm(out["G"]), np.linalg.norm(Gr), rtol=rt)
    assert out["G"].shape == tuple(ranks)
    for n in range(len(dims)):
        assert D.sin_max_angle(out["U"][n], Ur[n]) < at


def test_sparse_hooi_rank_limits_einval(ctx):
    rng = np.random.default_rng(3)
    dims = (100, 30, 20)
    idx, vals = coo(rng, dims, 800)
    sp = dt.SpTensor.from_host(ctx, dims, idx, vals.astype(np.float32))
    try:
        # above 64; above I; above the product of the other ranks; accepted
        for ranks in ((65, 8, 8), (8, 31, 8), (13, 3, 4), (20, 20, 12)):
            if ranks == (20, 20, 12):
                sp.hooi(max_iter=1, ranks=ranks)  # accepted
                continue
            with pytest.raises(dt.DTError) as e:
                sp.hooi(max_iter=1, ranks=ranks)
            assert e.value.code == dt.DT_EINVAL
    finally:
        sp.close()


@pytest.mark.parametrize("dims,nnz,dup", [((50, 40, 30), 5000, True), ((20, 10, 8, 6), 3000, True),
                                          ((1 << 20, 1 << 18, 1 << 16), 200000, False), ((7, 5), 400, True)])
def test_merge_duplicates_vs_oracle(ctx, dims, nnz, dup):
    """Coordinates merged at their first occurrence with FP64 sums in input
    order, survivors in input order; bit-exact against the oracle.  Without
    duplicates the list is unchanged."""
    rng = np.random.default_rng(nnz)
    if dup:
        idx = [rng.integers(0, d, nnz) for d in dims]
    else:
        # distinct coordinates: a random injective draw of composite keys
        keys = rng.choice(np.prod([min(d, 1 << 12) for d in dims], dtype=np.int64), nnz, replace=False)
        idx = []
        for d in dims:
            m = min(d, 1 << 12)
            idx.append(keys % m)
            keys = keys // m
    vals = rng.standard_normal(nnz)
    gi, gv = dt.sp_merge_duplicates(ctx, dims, idx, vals)
    ri, rv = S.merge_first_occurrence(idx, vals)
    assert len(gv) == len(rv)
    for a, b in zip(gi, ri):
        assert np.array_equal(a, b)
    assert np.array_equal(gv, rv)
    if not dup:
        assert len(gv) == nnz and np.array_equal(gv, vals)


def test_merge_duplicates_rejects_out_of_range(ctx):
    idx = [np.array([0, 5, 2]), np.array([1, 1, 9])]
    with pytest.raises(dt.DTError) as e:
        dt.sp_merge_duplicates(ctx, (6, 9), idx, np.ones(3))
    assert e.value.code == dt.DT_EINVAL
