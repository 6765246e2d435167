"""dt_gram_noalloc (the Gram without device allocations, scratch in a static
device array) against dt_gram through the C ABI: two calls in a row, the
second on a smaller input, for every input type; the oversize error."""
import ctypes as C

import numpy as np
import pytest

import dtucker as dt

pytestmark = pytest.mark.gpu


def _noalloc(ctx, A, n, dtS):
    L = dt.lib()
    f = L.dt_gram_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_size_t), C.c_int, C.c_void_p,
                  C.c_int]
    f.restype = C.c_int
    I = A.shape[n]
    S = ctx.array((I, I), dtS)
    shp = (C.c_size_t * len(A.shape))(*A.shape)
    st = f(ctx.handle, A.ptr, A.dtype, len(A.shape), shp, n, S.ptr, dtS)
    return st, S


def _err(ctx):
    return ctx.L.dt_last_error(ctx.handle).decode()


@pytest.mark.parametrize("dtA", [dt.DT_F32, dt.DT_BF16, dt.DT_F64])
@pytest.mark.parametrize("dtS", [dt.DT_F64, dt.DT_F32])
def test_gram_noalloc_matches_gram_twice(ctx, dtA, dtS):
    rng = np.random.default_rng(7 + dtA + 10 * dtS)
    # the second input is smaller, so a stale scratch from the first would show
    for shape, n in (((96, 80, 72), 1), ((40, 30, 20), 0)):
        print("gram_noalloc case", shape, n, "dtA", dtA, "dtS", dtS, flush=True)
        X = rng.standard_normal(shape)
        if dtA == dt.DT_F64:
            A = ctx.to_device(X)
        elif dtA == dt.DT_F32:
            A = ctx.to_device(X.astype(np.float32))
        else:
            A = ctx.to_device(dt.bf16_from_f32(X.astype(np.float32)))
        want = ctx.gram(A, n, dtS).download()
        st, S = _noalloc(ctx, A, n, dtS)
        assert st == dt.DT_OK, _err(ctx)
        got = S.download()
        # same kernels and the same split-K plan at these sizes: bit-identical
        assert np.array_equal(got, want)


def test_gram_noalloc_scratch_bound_split(ctx):
    """Order 1024 with an F32 S: the FP64 Gram (8 MiB) and two splits of FP32
    partials (4 MiB each) fill the 16 MiB scratch, so the no-alloc path runs
    two splits where dt_gram runs three; the partial sums add in another
    order, so the two agree at the FP32 accumulation level (and with FP64)."""
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1024, 2000)).astype(np.float32)
    A = ctx.to_device(X)
    want = ctx.gram(A, 0, dt.DT_F32).download().astype(np.float64)
    st, S = _noalloc(ctx, A, 0, dt.DT_F32)
    assert st == dt.DT_OK, _err(ctx)
    got = S.download().astype(np.float64)
    ref = X.astype(np.float64) @ X.astype(np.float64).T
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() / scale < 1e-5
    assert np.abs(got - want).max() / scale < 1e-5


def test_gram_noalloc_largest_order(ctx):
    """Order 2048 from FP32 data with an F64 S: one split of partials fills the
    scratch exactly; dt_gram also runs one split there, so bit-identical."""
    rng = np.random.default_rng(4)
    A = ctx.to_device(rng.standard_normal((2048, 300)).astype(np.float32))
    want = ctx.gram(A, 0, dt.DT_F64).download()
    st, S = _noalloc(ctx, A, 0, dt.DT_F64)
    assert st == dt.DT_OK, _err(ctx)
    assert np.array_equal(S.download(), want)


def test_gram_noalloc_above_the_scratch_is_einval(ctx):
    A = ctx.to_device(np.ones((2049, 4), np.float32))
    st, _ = _noalloc(ctx, A, 0, dt.DT_F64)
    assert st == dt.DT_EINVAL
    assert "maximum" in _err(ctx)
