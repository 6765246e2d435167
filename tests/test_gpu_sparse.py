"""Sparse path on the GPU: Zipf COO generator and duplicate merging, the
stable per-mode sort (bit-exact against the reference constructor's counting
sort), the TTMc and the sparse HOOI, all against oracle/sparse.py."""
import numpy as np
import pytest

import dtucker as dt
from oracle import dense as D
from oracle import philox
from oracle import sparse as S

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("dims,ndraws,seed", [((500, 300, 200), 20000, 1),
                                              ((1 << 20, 1 << 18, 1 << 16), 200000, 20000905),
                                              ((64, 64, 64, 8), 30000, 5)])
def test_zipf_generator_bit_exact(ctx, dims, ndraws, seed):
    """Indices, merge order and merged values identical to the oracle."""
    print("zipf", dims, ndraws, flush=True)
    idx, val, nnz = dt.sp_gen_zipf(ctx, dims, ndraws, 1.1, seed)
    ridx, rval = S.gen_zipf(dims, ndraws, 1.1, seed)
    assert nnz == len(rval)
    for n in range(len(dims)):
        assert np.array_equal(idx[n].download()[:nnz], ridx[n])
    assert np.array_equal(val.download()[:nnz], rval)


@pytest.mark.parametrize("n,dim,seed", [(1000, 7, 1), (100000, 1000, 2), (3000000, 1 << 20, 3),
                                         (1, 5, 4), (4095, 3, 5), (4096, 300, 6), (4097, 1 << 17, 7),
                                         (65537, 1 << 24, 8), (20000, 1, 9), (12289, 1 << 25, 10)])
def test_sort_mode_bit_exact(ctx, n, dim, seed):
    """perm == argsort(kind=stable) (the constructor's counting sort), row pointers
    == histogram scan; sizes around the sort's 4096-item tiles, one to four
    8-bit passes, a single key value."""
    print("sort", n, dim, flush=True)
    rng = np.random.default_rng(seed)
    if dim >= 1 << 24:
        keys = rng.integers(0, dim, n).astype(np.int32)
    else:
        keys = np.minimum((rng.pareto(1.1, n) * 3).astype(np.int64), dim - 1).astype(np.int32)
    perm, rowptr = dt.sp_sort_mode(ctx, ctx.to_device(keys), n, dim)
    rp, rr = S.sort_mode(keys, dim)
    assert np.array_equal(perm.download(), rp)
    assert np.array_equal(rowptr.download(), rr)


def make_coo(rng, dims, nnz, long_row=None):
    """Skewed random COO (duplicates merged) with empty rows and, optionally, one
    mode-0 row longer than a TTMc chunk."""
    idx = [np.minimum((rng.pareto(1.0, nnz) * d / 8).astype(np.int64), d - 1) for d in dims]
    if long_row is not None:
        idx[0][:long_row] = 3
    vals = rng.random(nnz).astype(np.float32)
    return S.merge_duplicates(dims, idx, vals)


def test_sp_create_modes_and_slice_mode(ctx):
    rng = np.random.default_rng(7)
    dims = (120, 300, 300)
    idx, vals = make_coo(rng, dims, 5000)
    sp = dt.SpTensor.from_host(ctx, dims, idx, vals)
    try:
        assert sp.slice_mode == 1  # largest mode, first on ties
        for n in range(3):
            perm, rowptr, nrows = sp.mode_arrays(n)
            rp, rr = S.sort_mode(idx[n], dims[n])
            assert np.array_equal(perm, rp) and np.array_equal(rowptr, rr)
            assert nrows == int(np.count_nonzero(np.diff(rr)))
        assert np.isclose(sp.norm, np.sqrt(np.sum(vals.astype(np.float64) ** 2)), rtol=1e-12)
    finally:
        sp.close()


@pytest.mark.parametrize("dims,nnz,long_row", [((300, 200, 100), 40000, 9000), ((64, 4096, 512), 20000, None),
                                               ((2000, 50, 40), 3000, 2500)])
def test_sp_ttmc_vs_oracle(ctx, dims, nnz, long_row):
    """Every mode's TTMc (rows across chunk boundaries, empty rows) against the
    oracle's Kronecker accumulation; FP32 accumulation: 2e-6 relative."""
    print("sp_ttmc", dims, nnz, long_row, flush=True)
    rng = np.random.default_rng(sum(dims))
    idx, vals = make_coo(rng, dims, nnz, long_row)
    Us = [rng.standard_normal((d, 16)).astype(np.float32) for d in dims]
    sp = dt.SpTensor.from_host(ctx, dims, idx, vals)
    try:
        dU = [dt.upload_rows(ctx, u) for u in Us]
        for n in range(3):
            Y = dt.download_rows(sp.ttmc(n, dU))
            Yr = S.ttmc(dims, idx, vals, [u.astype(np.float64) for u in Us], n)
            assert rel(Y, Yr) < 2e-6, n
    finally:
        sp.close()


@pytest.mark.parametrize("case", ["one", "row63", "row64", "row65", "rows130", "two_rows_129", "fibres",
                                  "row65_wide", "fibres_wide"])
def test_sp_ttmc_tiny_edges(ctx, case):
    """Batch-boundary edges of the TTMc (batches of 64 records): a single
    nonzero, one row of 63 / 64 / 65 nonzeros, 130 rows of one nonzero each,
    two rows whose boundary falls inside the second batch, and two rows whose
    nonzeros share one index of a gathered mode in runs of 1, 8, 9, 17 and 64.
    The "_wide" cases put the same rows among gathered modes of thousands of
    rows (the runs' indices between 600 heavily used others)."""
    rng = np.random.default_rng(5)
    if case == "fibres_wide":
        # row 0 of mode 0: 600 mode-2 indices used 70 times each; rows 1 and
        # 3 as in "fibres" on three other mode-2 indices
        dims = (4, 20000, 3000)
        i0, i1, i2 = [], [], []
        for k2 in range(600):
            sel = rng.permutation(20000)[:70]
            i0 += [0] * 70
            i1 += list(sel)
            i2 += [k2] * 70
        for r, lens in ((1, (9, 17, 64)), (3, (1, 8, 0))):
            for j, ln in enumerate(lens):
                sel = rng.permutation(20000)[:ln]
                i0 += [r] * ln
                i1 += list(sel)
                i2 += [1000 + 997 * j] * ln
        idx = [np.array(i0), np.array(i1), np.array(i2)]
    elif case == "row65_wide":
        dims = (4, 5000, 4000)
        flat = rng.permutation(5000 * 4000)[:65]
        idx = [np.full(65, 2), flat % 5000, flat // 5000]
    elif case == "fibres":
        # row 1 of mode 0 holds runs of 9, 17 and 64 nonzeros sharing their
        # index in mode 2, row 3 runs of 1 and 8
        dims = (4, 64, 3)
        i0, i1, i2 = [], [], []
        for r, lens in ((1, (9, 17, 64)), (3, (1, 8, 0))):
            for k2, ln in enumerate(lens):
                sel = rng.permutation(64)[:ln]
                i0 += [r] * ln
                i1 += list(sel)
                i2 += [k2] * ln
        idx = [np.array(i0), np.array(i1), np.array(i2)]
    elif case == "one":
        dims = (5, 4, 3)
        idx = [np.array([3]), np.array([1]), np.array([2])]
    elif case.startswith("row"):
        k = int(case[3:]) if case != "rows130" else 130
        if case == "rows130":
            dims = (200, 40, 30)
            idx = [rng.permutation(200)[:k], rng.integers(0, 40, k), rng.integers(0, 30, k)]
        else:
            dims = (4, 50, 40)
            flat = rng.permutation(50 * 40)[:k]
            idx = [np.full(k, 2), flat % 50, flat // 50]
    else:
        dims = (3, 60, 50)
        flat = rng.permutation(60 * 50)[:129]
        idx = [np.where(np.arange(129) < 70, 0, 2), flat % 60, flat // 60]
    idx = [np.asarray(i, np.int64) for i in idx]
    vals = rng.random(len(idx[0])).astype(np.float32) + np.float32(0.5)
    idx, vals = S.merge_duplicates(dims, idx, vals)
    print("tiny", case, dims, len(vals), flush=True)
    Us = [rng.standard_normal((d, 16)).astype(np.float32) for d in dims]
    sp = dt.SpTensor.from_host(ctx, dims, idx, vals)
    try:
        dU = [dt.upload_rows(ctx, u) for u in Us]
        for n in range(3):
            Y = dt.download_rows(sp.ttmc(n, dU))
            Yr = S.ttmc(dims, idx, vals, [u.astype(np.float64) for u in Us], n)
            assert rel(Y, Yr) < 2e-6, (case, n)
    finally:
        sp.close()


def test_sparse_hooi_tiny_ranks(ctx):
    """A 20 x 18 x 17 tensor of 200 nonzeros at ranks (3, 4, 5) from a given start."""
    rng = np.random.default_rng(21)
    dims, ranks = (20, 18, 17), (3, 4, 5)
    idx, vals = make_coo(rng, dims, 200)
    U0 = [D.orth_posdiag(rng.standard_normal((d, r))) for d, r in zip(dims, ranks)]
    sp = dt.SpTensor.from_host(ctx, dims, idx, vals)
    try:
        out = sp.hooi(max_iter=3, ranks=ranks, U_init=U0)
    finally:
        sp.close()
    Ur, Gr, res = S.hooi(dims, idx, vals, U0, 3)
    print("tiny ranks residuals", out["residuals"], res, flush=True)
    assert np.allclose(out["residuals"], res, rtol=1e-5)
    for n in range(3):
        assert D.sin_max_angle(out["U"][n], Ur[n]) < 1e-3


def test_sparse_hooi_vs_oracle(ctx):
    """Three sweeps from the same start (Philox normals, positive-diagonal QR):
    residuals to 1e-5 relative, factors by principal angles."""
    rng = np.random.default_rng(11)
    dims = (400, 300, 200)
    idx, vals = make_coo(rng, dims, 60000)
    seed = 77
    sp = dt.SpTensor.from_host(ctx, dims, idx, vals)
    try:
        out = sp.hooi(max_iter=3, seed=seed)
    finally:
        sp.close()
    U0 = [D.orth_posdiag(philox.normal(d * 16, seed, 200 + m).reshape((d, 16), order="F"))
          for m, d in enumerate(dims)]
    Ur, Gr, res = S.hooi(dims, idx, vals, U0, 3)
    print("residuals gpu", out["residuals"], "oracle", res, flush=True)
    assert np.allclose(out["residuals"], res, rtol=1e-5)
    for n in range(3):
        assert D.sin_max_angle(out["U"][n], Ur[n]) < 1e-3
    G = out["G"]
    assert np.isclose(np.linalg.norm(G), np.linalg.norm(Gr), rtol=1e-5)


def test_sparse_hooi_ranks_below_16_from_given_start(ctx):
    """Ranks (10, 7, 12) (the reference's default truncation is 10) from the
    caller's start factors, against the oracle."""
    rng = np.random.default_rng(12)
    dims, ranks = (300, 250, 200), (10, 7, 12)
    idx, vals = make_coo(rng, dims, 40000)
    U0 = [D.orth_posdiag(rng.standard_normal((d, r))) for d, r in zip(dims, ranks)]
    sp = dt.SpTensor.from_host(ctx, dims, idx, vals)
    try:
        out = sp.hooi(max_iter=3, ranks=ranks, U_init=U0)
    finally:
        sp.close()
    Ur, Gr, res = S.hooi(dims, idx, vals, U0, 3)
    print("ranks", ranks, "residuals", out["residuals"], res, flush=True)
    assert np.allclose(out["residuals"], res, rtol=1e-5)
    assert out["G"].shape == ranks
    for n in range(3):
        assert D.sin_max_angle(out["U"][n], Ur[n]) < 1e-3


def test_sparse_bad_ranks_einval(ctx):
    """A rank above its mode's size is refused (ranks above 16 take the general path)."""
    idx, vals = make_coo(np.random.default_rng(1), (40, 30, 20), 500)
    sp = dt.SpTensor.from_host(ctx, (40, 30, 20), idx, vals)
    try:
        with pytest.raises(dt.DTError) as e:
            ctx.check(ctx.L.dt_hooi_sparse(ctx.handle, sp.handle, (dt.C.c_size_t * 3)(41, 8, 8), 1, 0, None, None,
                                           None, None, dt.C.byref(dt.C.c_double()), None))
        assert e.value.code == dt.DT_EINVAL
    finally:
        sp.close()


def test_sp_create_rejects_out_of_range_index(ctx):
    """An index outside its mode is refused at creation (the TTMc gathers by it)."""
    idx, vals = make_coo(np.random.default_rng(2), (40, 30, 20), 500)
    bad = [i.copy() for i in idx]
    bad[1][7] = 30
    with pytest.raises(dt.DTError) as e:
        dt.SpTensor.from_host(ctx, (40, 30, 20), bad, vals)
    assert e.value.code == dt.DT_EINVAL
    bad[1][7] = -1
    with pytest.raises(dt.DTError):
        dt.SpTensor.from_host(ctx, (40, 30, 20), bad, vals)
    sp = dt.SpTensor.from_host(ctx, (40, 30, 20), idx, vals)  # the context still works
    sp.close()


def test_sparse_full_size_generator_and_sort(ctx):
    """The sparse workload at its real size: 2^20 x 2^18 x 2^16, 1e8 Zipf(1.1)
    draws.  Size-independent checks: ~65.6 % of the draws survive the merge,
    merged keys strictly increase, values stay in (0, #draws of the key], the
    slice-mode permutation is the stable argsort (keys non-decreasing along it
    and, within equal keys, positions increasing) and the row pointers count it."""
    dims = (1 << 20, 1 << 18, 1 << 16)
    print("sparse 2^20x2^18x2^16 generate", flush=True)
    idx, val, nnz = dt.sp_gen_zipf(ctx, dims, 100_000_000, 1.1, 20000905)
    print("sparse 2^20x2^18x2^16 nnz", nnz, flush=True)
    assert 0.60 < nnz / 1e8 < 0.71
    i0 = idx[0].download()[:nnz].astype(np.int64)
    i1 = idx[1].download()[:nnz].astype(np.int64)
    i2 = idx[2].download()[:nnz].astype(np.int64)
    key = i0 | (i1 << 20) | (i2 << 38)
    assert np.all(np.diff(key) > 0)
    v = val.download()[:nnz]
    assert np.all(v > 0)
    sp = dt.SpTensor(ctx, dims, idx, val, nnz)
    try:
        assert sp.slice_mode == 0
        perm, rowptr, nrows = sp.mode_arrays(0)
        ks = i0[perm]
        assert np.all(np.diff(ks) >= 0)
        same = np.diff(ks) == 0
        assert np.all(np.diff(perm.astype(np.int64))[same] > 0)
        assert np.array_equal(np.diff(rowptr.astype(np.int64)), np.bincount(i0, minlength=dims[0]))
        print("sparse 2^20x2^18x2^16 mode-0 rows touched", nrows, flush=True)
    finally:
        sp.close()


def test_sparse_full_size_ttmc_rows(ctx):
    """The TTMc at the sparse workload's real size (2^20 x 2^18 x 2^16, 1e8
    Zipf(1.1) draws, 65.6M nonzeros) for every mode: the 16 longest rows (up to
    4.2M nonzeros each, split over many ranges and summed by the fix-up) and 2000
    random rows, each recomputed on the CPU in FP64 as the matrix product
    (val * U_a[i_a])^T U_b[i_b] over the row's nonzeros (entry r_a + 16 r_b),
    which is the Kronecker sum written as one GEMM.  FP32 accumulation: 1e-5
    relative per row (Frobenius).  Every row of Y is finite and the rows with no
    nonzero are zero."""
    dims = (1 << 20, 1 << 18, 1 << 16)
    print("sparse 2^20x2^18x2^16 ttmc: generate", flush=True)
    idx, val, nnz = dt.sp_gen_zipf(ctx, dims, 100_000_000, 1.1, 20000905)
    hidx = [idx[m].download()[:nnz].astype(np.int64) for m in range(3)]
    v = val.download()[:nnz].astype(np.float64)
    sp = dt.SpTensor(ctx, dims, idx, val, nnz)
    try:
        rng = np.random.default_rng(99)
        Us = [rng.standard_normal((d, 16)).astype(np.float32) for d in dims]
        dU = [dt.upload_rows(ctx, u) for u in Us]
        U64 = [u.astype(np.float64) for u in Us]
        for n in range(3):
            print("sparse 2^20x2^18x2^16 ttmc mode", n, flush=True)
            Y = sp.ttmc(n, dU)
            Yh = dt.download_rows(Y)
            Y.free()
            a, b = [m for m in range(3) if m != n]
            order = np.argsort(hidx[n], kind="stable")
            cnt = np.bincount(hidx[n], minlength=dims[n])
            rp = np.concatenate(([0], np.cumsum(cnt)))
            assert np.all(np.isfinite(Yh))
            assert not np.any(Yh[cnt == 0])
            nonempty = np.flatnonzero(cnt)
            rows = np.unique(np.concatenate((np.argsort(cnt)[-16:], rng.choice(nonempty, 2000, replace=False))))
            worst = 0.0
            for r in rows:
                sl = order[rp[r]:rp[r + 1]]
                M = np.zeros((16, 16))
                for c0 in range(0, len(sl), 1 << 20):
                    s = sl[c0:c0 + (1 << 20)]
                    M += (U64[a][hidx[a][s]] * v[s, None]).T @ U64[b][hidx[b][s]]
                ref = M.ravel(order="F")
                e = float(np.linalg.norm(Yh[r] - ref) / max(np.linalg.norm(ref), 1e-300))
                worst = max(worst, e)
                assert e < 1e-5, (n, int(r), int(cnt[r]), e)
            print("sparse 2^20x2^18x2^16 ttmc mode", n, "rows checked", len(rows), "worst rel", worst, flush=True)
            del Yh
    finally:
        sp.close()
