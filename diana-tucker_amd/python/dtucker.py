"""ctypes binding of libdiana_hip.so (the C ABI declared in include/diana_hip.h).

Used by the test-suite and bench.py to drive the HIP library through exactly the
entry points a foreign host would bind.  Nothing here computes: every numerical
call goes to the shared library, and loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "libdiana_hip.so")

DT_OK, DT_EINVAL, DT_ENOMEM, DT_EHIP, DT_ENUMERIC, DT_EUNSUPPORTED = range(6)
DT_F64, DT_F32, DT_BF16, DT_I32 = range(4)
DT_MAX_NDIM = 8
INIT_GIVEN, INIT_QR, INIT_HOSVD, INIT_RANDOM, INIT_LEARNED = range(5)
SCHED_REFERENCE, SCHED_TREE = 0, 1
FLAG_DIRECT_RESIDUAL, FLAG_FP32_POLISH, FLAG_FP64_FIT, FLAG_REFERENCE_ALS = 1, 2, 4, 8
FLAG_LEARNED_CHECK, FLAG_SKETCH_KRON, FLAG_LEARNED_ACCEPT = 16, 32, 64
LEARNED_ACCEPT_FIT = 1e-4
SKETCH_GAUSSIAN, SKETCH_KRONECKER = 0, 1

_NP_OF = {DT_F64: np.float64, DT_F32: np.float32, DT_BF16: np.uint16, DT_I32: np.int32}
_SIZE_OF = {DT_F64: 8, DT_F32: 4, DT_BF16: 2, DT_I32: 4}


class DTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"diana_hip error {code}: {msg}")
        self.code = code


class TuckerSpec(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("shape", C.c_size_t * DT_MAX_NDIM),
                ("ranks", C.c_size_t * DT_MAX_NDIM), ("noise", C.c_double),
                ("decay", C.c_double), ("seed", C.c_uint64)]


class HooiArgs(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("dtype", C.c_int32), ("A", C.c_void_p),
                ("A_f32", C.c_void_p), ("shape", C.c_size_t * DT_MAX_NDIM),
                ("ranks", C.c_size_t * DT_MAX_NDIM), ("max_iter", C.c_int32),
                ("inner_iter", C.c_int32), ("init", C.c_int32), ("schedule", C.c_int32),
                ("tol", C.c_double), ("U_init", C.c_void_p * DT_MAX_NDIM),
                ("seed", C.c_uint64), ("flags", C.c_int32), ("model", C.c_void_p)]


class HooiResult(C.Structure):
    _fields_ = [("U", C.c_void_p * DT_MAX_NDIM), ("G", C.c_void_p),
                ("residual_per_sweep", C.POINTER(C.c_double)), ("sweeps", C.c_int32),
                ("init_used", C.c_int32), ("A_norm", C.c_double), ("G_norm", C.c_double),
                ("residual", C.c_double), ("residual_direct", C.c_double),
                ("init_ms", C.c_double), ("sweeps_ms", C.c_double),
                ("learned_first_res", C.c_double), ("hosvd_first_res", C.c_double)]


_lib = None


def lib():
    """Load the shared library once; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `make -C diana-tucker_amd` "
                "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, sz, i32, dbl, u64 = C.c_void_p, C.c_size_t, C.c_int, C.c_double, C.c_uint64
        sig = {
            "dt_create": [C.POINTER(vp), i32], "dt_destroy": [vp], "dt_last_error": [vp],
            "dt_malloc": [vp, C.POINTER(vp), sz], "dt_free": [vp, vp],
            "dt_h2d": [vp, vp, vp, sz], "dt_d2h": [vp, vp, vp, sz], "dt_d2d": [vp, vp, vp, sz],
            "dt_zero": [vp, vp, sz], "dt_sync": [vp],
            "dt_convert": [vp, vp, i32, vp, i32, sz],
            "dt_timing_enable": [vp, i32], "dt_timing_reset": [vp], "dt_timing_filter": [vp, C.c_char_p],
            "dt_timing_get": [vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(dbl)],
            "dt_timing_list": [vp, C.c_char_p, sz],
            "dt_timer_start": [vp], "dt_timer_stop": [vp, C.POINTER(dbl)],
            "dt_ttm": [vp, vp, i32, i32, C.POINTER(sz), vp, i32, sz, i32, vp, i32],
            "dt_ttmc": [vp, vp, i32, i32, C.POINTER(sz), C.POINTER(vp), i32, C.POINTER(sz), C.POINTER(i32),
                        i32, vp, i32],
            "dt_gram": [vp, vp, i32, i32, C.POINTER(sz), i32, vp, i32],
            "dt_gram_noalloc": [vp, vp, i32, i32, C.POINTER(sz), i32, vp, i32],
            "dt_ttt_except": [vp, vp, vp, i32, i32, C.POINTER(sz), sz, i32, vp],
            "dt_sumsq": [vp, vp, i32, sz, C.POINTER(dbl)],
            "dt_gemm_f64": [vp, i32, i32, sz, sz, sz, dbl, vp, sz, vp, sz, dbl, vp, sz],
            "dt_qr_f64": [vp, vp, vp, vp, sz, sz],
            "dt_inverse_f64": [vp, vp, vp, sz],
            "dt_syev_top_f64": [vp, vp, sz, sz, vp, vp],
            "dt_orth_rows_f64": [vp, vp, sz, sz],
            "dt_gen_normal": [vp, vp, i32, sz, u64, u64],
            "dt_gen_uniform": [vp, vp, i32, sz, u64, u64],
            "dt_gen_tucker": [vp, C.POINTER(TuckerSpec), vp, vp, C.POINTER(vp),
                              C.POINTER(dbl), C.POINTER(dbl)],
            "dt_hooi_dense": [vp, C.POINTER(HooiArgs), C.POINTER(HooiResult)],
            "dt_hooi_create": [vp, C.POINTER(HooiArgs), C.POINTER(vp)],
            "dt_hooi_sweep": [vp, i32, C.POINTER(dbl)],
            "dt_hooi_get": [vp, C.POINTER(HooiResult)],
            "dt_hooi_fit": [vp, i32, C.POINTER(dbl)],
            "dt_hooi_destroy": [vp],
            "dt_mlp_create": [vp, C.POINTER(vp), i32, i32, i32, u64],
            "dt_mlp_destroy": [vp, vp],
            "dt_mlp_shape": [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(C.c_int64)],
            "dt_mlp_get_params": [vp, vp, C.POINTER(C.c_float), C.c_int64],
            "dt_mlp_set_params": [vp, vp, C.POINTER(C.c_float), C.c_int64],
            "dt_mlp_forward": [vp, vp, vp, C.c_int64, vp],
            "dt_mlp_loss_grad": [vp, vp, vp, vp, C.c_int64, C.POINTER(dbl), C.POINTER(C.c_float)],
            "dt_mlp_adam_step": [vp, vp, C.c_float],
            "dt_sketch_features": [vp, vp, i32, i32, C.POINTER(sz), i32, i32, u64, vp],
            "dt_sketch_features_kron": [vp, vp, i32, i32, C.POINTER(sz), i32, i32, u64, vp],
            "dt_learned_start": [vp, vp, vp, i32, i32, C.POINTER(sz), C.POINTER(sz), i32, u64, C.POINTER(vp)],
            "dt_sp_gen_zipf": [vp, i32, C.POINTER(sz), C.c_int64, dbl, u64, C.POINTER(vp), vp,
                               C.POINTER(C.c_int64)],
            "dt_sp_sort_mode": [vp, C.c_int64, vp, sz, vp, vp],
            "dt_sp_create": [vp, i32, C.POINTER(sz), C.c_int64, C.POINTER(vp), vp, C.POINTER(vp)],
            "dt_sp_destroy": [vp],
            "dt_sp_info": [vp, C.POINTER(i32), C.POINTER(C.c_int64), C.POINTER(dbl)],
            "dt_sp_mode_arrays": [vp, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_int64)],
            "dt_sp_ttmc": [vp, vp, i32, C.POINTER(vp), C.POINTER(sz), vp],
            "dt_hooi_sparse": [vp, vp, C.POINTER(sz), i32, u64, C.POINTER(vp), C.POINTER(vp), vp,
                               C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)],
            "dt_sp_create_f64": [vp, i32, C.POINTER(sz), C.c_int64, C.POINTER(vp), vp, C.POINTER(vp)],
            "dt_sp_merge_duplicates": [vp, i32, C.POINTER(sz), C.c_int64, C.POINTER(vp), vp, C.POINTER(C.c_int64)],
            "dt_sp_ttmc_mtr": [vp, vp, i32, C.POINTER(vp), C.POINTER(sz), vp],
            "dt_sp_ttmc_mtl": [vp, vp, i32, C.POINTER(vp), C.POINTER(sz), vp, vp],
            "dt_sp_hooi_create": [vp, vp, C.POINTER(sz), u64, C.POINTER(vp), C.POINTER(vp)],
            "dt_sp_hooi_sweep": [vp, C.POINTER(dbl), C.POINTER(dbl)],
            "dt_sp_hooi_get": [vp, C.POINTER(vp), vp, C.POINTER(dbl)],
            "dt_sp_hooi_destroy": [vp],
        }
        for name, argtypes in sig.items():
            f = getattr(L, name)
            f.argtypes = argtypes
            f.restype = C.c_char_p if name == "dt_last_error" else C.c_int
        _lib = L
    return _lib


def exported_symbols():
    """Names this binding expects the library to export."""
    return ["dt_create", "dt_destroy", "dt_last_error", "dt_malloc", "dt_free", "dt_h2d",
            "dt_d2h", "dt_d2d", "dt_zero", "dt_sync", "dt_convert", "dt_timing_enable",
            "dt_timing_reset", "dt_timing_filter", "dt_timing_get", "dt_timing_list", "dt_timer_start",
            "dt_timer_stop", "dt_ttm", "dt_ttmc", "dt_gram", "dt_gram_noalloc", "dt_ttt_except", "dt_sumsq", "dt_gemm_f64",
            "dt_qr_f64", "dt_inverse_f64", "dt_syev_top_f64", "dt_orth_rows_f64",
            "dt_gen_normal", "dt_gen_uniform", "dt_gen_tucker", "dt_hooi_dense",
            "dt_hooi_create", "dt_hooi_sweep", "dt_hooi_fit", "dt_hooi_get", "dt_hooi_destroy",
            "dt_mlp_create", "dt_mlp_destroy", "dt_mlp_shape", "dt_mlp_get_params", "dt_mlp_set_params",
            "dt_mlp_forward", "dt_mlp_loss_grad", "dt_mlp_adam_step", "dt_sketch_features",
            "dt_sketch_features_kron", "dt_learned_start", "dt_sp_gen_zipf", "dt_sp_sort_mode",
            "dt_sp_create", "dt_sp_destroy", "dt_sp_info", "dt_sp_mode_arrays", "dt_sp_ttmc",
            "dt_hooi_sparse", "dt_sp_create_f64", "dt_sp_merge_duplicates", "dt_sp_ttmc_mtr", "dt_sp_ttmc_mtl",
            "dt_sp_hooi_create", "dt_sp_hooi_sweep", "dt_sp_hooi_get", "dt_sp_hooi_destroy"]


class DeviceArray:
    """A device allocation with a numpy-like shape (column-major semantics)."""

    def __init__(self, ctx: "Context", shape, dtype: int):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = dtype
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * _SIZE_OF[dtype]
        self.ptr = ctx.malloc(self.nbytes)

    def upload(self, host: np.ndarray):
        h = np.asarray(host)
        want = _NP_OF[self.dtype]
        if h.dtype != want:
            h = h.astype(want)
        flat = np.ascontiguousarray(h.reshape(-1, order="F"))
        assert flat.size == self.size, (flat.size, self.size)
        self.ctx.h2d(self.ptr, flat.ctypes.data, self.nbytes)
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.size, dtype=_NP_OF[self.dtype])
        self.ctx.d2h(out.ctypes.data, self.ptr, self.nbytes)
        return out.reshape(self.shape, order="F")

    def free(self):
        if self.ptr:
            self.ctx.free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                self.ctx.free(self.ptr)
        except Exception:
            pass


def bf16_from_f32(x: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even FP32 -> BF16 bit patterns (uint16)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16).astype(np.uint16)
    return r.reshape(np.shape(x))


def f32_from_bf16(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


class Context:
    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.dt_create(C.byref(h), device)
        if rc != DT_OK:
            raise DTError(rc, f"dt_create(device={device}) failed")
        self.handle = h

    def check(self, rc):
        if rc != DT_OK:
            msg = self.L.dt_last_error(self.handle)
            raise DTError(rc, msg.decode() if msg else "")

    def close(self):
        if self.handle:
            self.L.dt_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # memory
    def malloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self.check(self.L.dt_malloc(self.handle, C.byref(p), nbytes))
        return p.value

    def free(self, ptr):
        self.check(self.L.dt_free(self.handle, ptr))

    def h2d(self, dst, src, nbytes):
        self.check(self.L.dt_h2d(self.handle, dst, src, nbytes))

    def d2h(self, dst, src, nbytes):
        self.check(self.L.dt_d2h(self.handle, dst, src, nbytes))

    def sync(self):
        self.check(self.L.dt_sync(self.handle))

    def array(self, shape, dtype=DT_F32) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def zero(self, A: DeviceArray):
        self.check(self.L.dt_zero(self.handle, A.ptr, A.nbytes))

    def convert(self, dst: DeviceArray, src: DeviceArray):
        assert dst.size == src.size
        self.check(self.L.dt_convert(self.handle, dst.ptr, dst.dtype, src.ptr, src.dtype, src.size))

    def to_device(self, host: np.ndarray, dtype=None) -> DeviceArray:
        h = np.asarray(host)
        if dtype is None:
            dtype = {np.dtype(np.float64): DT_F64, np.dtype(np.float32): DT_F32,
                     np.dtype(np.uint16): DT_BF16, np.dtype(np.int32): DT_I32}[h.dtype]
        return DeviceArray(self, h.shape, dtype).upload(h)

    # timing
    def timing(self, on=True):
        self.check(self.L.dt_timing_enable(self.handle, 1 if on else 0))

    def timing_filter(self, prefix: str = ""):
        self.check(self.L.dt_timing_filter(self.handle, prefix.encode()))

    def timing_reset(self):
        self.check(self.L.dt_timing_reset(self.handle))

    def timing_get(self, name: str):
        cnt = C.c_int64()
        ms = C.c_double()
        self.check(self.L.dt_timing_get(self.handle, name.encode(), C.byref(cnt), C.byref(ms)))
        return cnt.value, ms.value

    def timing_list(self):
        buf = C.create_string_buffer(1 << 16)
        self.check(self.L.dt_timing_list(self.handle, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def timer_start(self):
        self.check(self.L.dt_timer_start(self.handle))

    def timer_stop(self) -> float:
        ms = C.c_double()
        self.check(self.L.dt_timer_stop(self.handle, C.byref(ms)))
        return ms.value

    # operators
    def ttm(self, A: DeviceArray, M: DeviceArray, n: int) -> DeviceArray:
        J = M.shape[0]
        assert M.shape[1] == A.shape[n]
        out_shape = list(A.shape)
        out_shape[n] = J
        ydt = DT_F64 if A.dtype == DT_F64 else DT_F32
        Y = self.array(out_shape, ydt)
        shp = (C.c_size_t * len(A.shape))(*A.shape)
        self.check(self.L.dt_ttm(self.handle, A.ptr, A.dtype, len(A.shape), shp, M.ptr, M.dtype,
                                 J, n, Y.ptr, ydt))
        return Y

    def ttmc(self, A: DeviceArray, Ms, idx) -> DeviceArray:
        """Y = A x_{idx[0]} Ms[0] x_{idx[1]} Ms[1] ... (all Ms of one dtype)."""
        assert len(Ms) == len(idx) and len(Ms) >= 1
        shape = list(A.shape)
        for M, n in zip(Ms, idx):
            if 0 <= n < len(shape):  # a bad mode is left for the library to reject
                assert M.shape[1] == shape[n] and M.dtype == Ms[0].dtype
                shape[n] = M.shape[0]
        ydt = DT_F64 if A.dtype == DT_F64 else DT_F32
        Y = self.array(shape, ydt)
        k = len(Ms)
        shp = (C.c_size_t * len(A.shape))(*A.shape)
        mp = (C.c_void_p * k)(*[M.ptr for M in Ms])
        js = (C.c_size_t * k)(*[M.shape[0] for M in Ms])
        ix = (C.c_int * k)(*idx)
        self.check(self.L.dt_ttmc(self.handle, A.ptr, A.dtype, len(A.shape), shp, mp, Ms[0].dtype, js, ix,
                                  k, Y.ptr, ydt))
        return Y

    def gram(self, A: DeviceArray, n: int, dtS=DT_F64) -> DeviceArray:
        I = A.shape[n]
        S = self.array((I, I), dtS)
        shp = (C.c_size_t * len(A.shape))(*A.shape)
        self.check(self.L.dt_gram(self.handle, A.ptr, A.dtype, len(A.shape), shp, n, S.ptr, dtS))
        return S

    def ttt_except(self, A: DeviceArray, B: DeviceArray, n: int) -> DeviceArray:
        assert A.dtype == B.dtype
        Cm = self.array((A.shape[n], B.shape[n]), DT_F64)
        shp = (C.c_size_t * len(A.shape))(*A.shape)
        self.check(self.L.dt_ttt_except(self.handle, A.ptr, B.ptr, A.dtype, len(A.shape), shp,
                                        B.shape[n], n, Cm.ptr))
        return Cm

    def sumsq(self, A: DeviceArray) -> float:
        out = C.c_double()
        self.check(self.L.dt_sumsq(self.handle, A.ptr, A.dtype, A.size, C.byref(out)))
        return out.value

    def gemm(self, tA, tB, A: DeviceArray, B: DeviceArray, alpha=1.0) -> DeviceArray:
        m = A.shape[1] if tA else A.shape[0]
        k = A.shape[0] if tA else A.shape[1]
        n = B.shape[0] if tB else B.shape[1]
        Cm = self.array((m, n), DT_F64)
        self.check(self.L.dt_gemm_f64(self.handle, int(tA), int(tB), m, n, k, alpha, A.ptr,
                                      A.shape[0], B.ptr, B.shape[0], 0.0, Cm.ptr, m))
        return Cm

    def qr(self, A: DeviceArray, want_r=True):
        m, n = A.shape
        Q = self.array((m, n), DT_F64)
        R = self.array((n, n), DT_F64) if want_r else None
        self.check(self.L.dt_qr_f64(self.handle, Q.ptr, R.ptr if R else None, A.ptr, m, n))
        return Q, R

    def inverse(self, A: DeviceArray) -> DeviceArray:
        n = A.shape[0]
        Cm = self.array((n, n), DT_F64)
        self.check(self.L.dt_inverse_f64(self.handle, Cm.ptr, A.ptr, n))
        return Cm

    def syev_top(self, S: DeviceArray, r: int):
        n = S.shape[0]
        V = self.array((n, r), DT_F64)
        w = self.array((r,), DT_F64)
        self.check(self.L.dt_syev_top_f64(self.handle, S.ptr, n, r, V.ptr, w.ptr))
        return V, w

    def orth_rows(self, A: DeviceArray):
        self.check(self.L.dt_orth_rows_f64(self.handle, A.ptr, A.shape[0], A.shape[1]))
        return A

    def gen_normal(self, n, dtype, seed, stream) -> DeviceArray:
        X = self.array((n,), dtype)
        self.check(self.L.dt_gen_normal(self.handle, X.ptr, dtype, n, seed, stream))
        return X

    def gen_uniform(self, n, dtype, seed, stream) -> DeviceArray:
        X = self.array((n,), dtype)
        self.check(self.L.dt_gen_uniform(self.handle, X.ptr, dtype, n, seed, stream))
        return X

    def gen_tucker(self, shape, ranks, noise, seed, decay=0.0, f32=True, bf16=False,
                   want_factors=False):
        spec = TuckerSpec()
        spec.ndim = len(shape)
        for i, (s, r) in enumerate(zip(shape, ranks)):
            spec.shape[i] = s
            spec.ranks[i] = r
        spec.noise = noise
        spec.decay = decay
        spec.seed = seed
        A = self.array(shape, DT_F32)
        Ab = self.array(shape, DT_BF16) if bf16 else None
        Us = [self.array((s, r), DT_F64) for s, r in zip(shape, ranks)] if want_factors else None
        uptr = (C.c_void_p * len(shape))(*[u.ptr for u in Us]) if Us else None
        nx, na = C.c_double(), C.c_double()
        self.check(self.L.dt_gen_tucker(self.handle, C.byref(spec), A.ptr, Ab.ptr if Ab else None,
                                        uptr, C.byref(nx), C.byref(na)))
        return {"A": A, "A_bf16": Ab, "U": Us, "normX": nx.value, "normA": na.value}

    # HOOI
    def _hooi_args(self, A, ranks, max_iter=10, inner_iter=5, init=INIT_HOSVD,
                   schedule=SCHED_TREE, tol=0.0, U_init=None, seed=0, flags=0, A_f32=None,
                   model=None):
        args = HooiArgs()
        args.ndim = len(A.shape)
        args.dtype = A.dtype
        args.A = A.ptr
        args.A_f32 = A_f32.ptr if A_f32 is not None else None
        for i, (s, r) in enumerate(zip(A.shape, ranks)):
            args.shape[i] = s
            args.ranks[i] = r
        args.max_iter = max_iter
        args.inner_iter = inner_iter
        args.init = init
        args.schedule = schedule
        args.tol = tol
        if U_init is not None:
            for i, u in enumerate(U_init):
                args.U_init[i] = u.ptr
        args.seed = seed
        args.flags = flags
        args.model = model
        return args

    # learned start
    def rows_array(self, rows, width) -> DeviceArray:
        """Row-major rows x width FP32 device buffer (stored as a width x rows column-major array)."""
        return self.array((width, rows), DT_F32)

    def sketch_features(self, A: DeviceArray, n: int, S: int, seed: int) -> np.ndarray:
        X = self.rows_array(A.shape[n], S)
        shp = (C.c_size_t * len(A.shape))(*A.shape)
        self.check(self.L.dt_sketch_features(self.handle, A.ptr, A.dtype, len(A.shape), shp, n, S, seed,
                                             X.ptr))
        return X.download().T.copy()

    def sketch_features_kron(self, A: DeviceArray, n: int, q: int, seed: int) -> np.ndarray:
        S = q ** (len(A.shape) - 1)
        X = self.rows_array(A.shape[n], S)
        shp = (C.c_size_t * len(A.shape))(*A.shape)
        self.check(self.L.dt_sketch_features_kron(self.handle, A.ptr, A.dtype, len(A.shape), shp, n, q, seed,
                                                  X.ptr))
        return X.download().T.copy()

    def learned_start(self, mlp: "MLP", A: DeviceArray, ranks, seed: int, sketch: int = SKETCH_GAUSSIAN):
        N = len(A.shape)
        Us = [self.array((A.shape[i], ranks[i]), DT_F64) for i in range(N)]
        shp = (C.c_size_t * N)(*A.shape)
        rk = (C.c_size_t * N)(*ranks)
        up = (C.c_void_p * N)(*[u.ptr for u in Us])
        self.check(self.L.dt_learned_start(self.handle, mlp.handle, A.ptr, A.dtype, N, shp, rk, sketch, seed, up))
        return Us

    def hooi_dense(self, A: DeviceArray, ranks, want_factors=True, **kw):
        args = self._hooi_args(A, ranks, **kw)
        N = len(A.shape)
        res = HooiResult()
        Us = [self.array((A.shape[i], ranks[i]), DT_F64) for i in range(N)] if want_factors else []
        G = self.array(tuple(ranks), DT_F64) if want_factors else None
        for i, u in enumerate(Us):
            res.U[i] = u.ptr
        res.G = G.ptr if G else None
        per = (C.c_double * max(1, args.max_iter))()
        res.residual_per_sweep = per
        self.check(self.L.dt_hooi_dense(self.handle, C.byref(args), C.byref(res)))
        out = {k: getattr(res, k) for k in ("sweeps", "init_used", "A_norm", "G_norm", "residual",
                                            "residual_direct", "init_ms", "sweeps_ms",
                                            "learned_first_res", "hosvd_first_res")}
        out["residuals"] = [per[i] for i in range(res.sweeps)]
        out["U"] = [u.download() for u in Us]
        out["G"] = G.download() if G else None
        return out

    def hooi_session(self, A: DeviceArray, ranks, **kw):
        return HooiSession(self, A, ranks, **kw)


class HooiSession:
    """dt_hooi_create / dt_hooi_sweep / dt_hooi_get: one sweep per call."""

    def __init__(self, ctx: Context, A: DeviceArray, ranks, **kw):
        self.ctx = ctx
        self.A = A
        self.ranks = list(ranks)
        self.args = ctx._hooi_args(A, ranks, **kw)
        h = C.c_void_p()
        ctx.check(ctx.L.dt_hooi_create(ctx.handle, C.byref(self.args), C.byref(h)))
        self.h = h

    def sweep(self, fp32=False) -> float:
        r = C.c_double()
        self.ctx.check(self.ctx.L.dt_hooi_sweep(self.h, 1 if fp32 else 0, C.byref(r)))
        return r.value

    def fit(self, fp64=False) -> float:
        """Residual of the current factors on A_f32 (or A), core formed again."""
        r = C.c_double()
        self.ctx.check(self.ctx.L.dt_hooi_fit(self.h, 1 if fp64 else 0, C.byref(r)))
        return r.value

    def result(self, want_factors=True):
        N = len(self.A.shape)
        res = HooiResult()
        Us = [self.ctx.array((self.A.shape[i], self.ranks[i]), DT_F64) for i in range(N)] \
            if want_factors else []
        G = self.ctx.array(tuple(self.ranks), DT_F64) if want_factors else None
        for i, u in enumerate(Us):
            res.U[i] = u.ptr
        res.G = G.ptr if G else None
        self.ctx.check(self.ctx.L.dt_hooi_get(self.h, C.byref(res)))
        out = {k: getattr(res, k) for k in ("sweeps", "init_used", "A_norm", "G_norm", "residual",
                                            "init_ms", "sweeps_ms")}
        out["U"] = [u.download() for u in Us]
        out["G"] = G.download() if G else None
        return out

    def close(self):
        if self.h:
            self.ctx.L.dt_hooi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MLP:
    """Row-wise 3-layer MLP of the learned start (dt_mlp_* entry points)."""

    def __init__(self, ctx: Context, d_in: int, hidden: int, d_out: int, seed: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(ctx.L.dt_mlp_create(ctx.handle, C.byref(h), d_in, hidden, d_out, seed))
        self.handle = h
        a, b, c, n = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        ctx.L.dt_mlp_shape(h, C.byref(a), C.byref(b), C.byref(c), C.byref(n))
        self.d_in, self.hidden, self.d_out, self.nparam = a.value, b.value, c.value, n.value

    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        self.ctx.check(self.ctx.L.dt_mlp_get_params(self.ctx.handle, self.handle,
                                                     out.ctypes.data_as(C.POINTER(C.c_float)), self.nparam))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self.ctx.check(self.ctx.L.dt_mlp_set_params(self.ctx.handle, self.handle,
                                                     p.ctypes.data_as(C.POINTER(C.c_float)), self.nparam))

    def forward(self, X: DeviceArray, rows: int) -> DeviceArray:
        Z = self.ctx.rows_array(rows, self.d_out)
        self.ctx.check(self.ctx.L.dt_mlp_forward(self.ctx.handle, self.handle, X.ptr, rows, Z.ptr))
        return Z

    def loss_grad(self, X: DeviceArray, T: DeviceArray, rows: int, want_grad=False):
        loss = C.c_double()
        g = np.empty(self.nparam, dtype=np.float32) if want_grad else None
        gp = g.ctypes.data_as(C.POINTER(C.c_float)) if want_grad else None
        self.ctx.check(self.ctx.L.dt_mlp_loss_grad(self.ctx.handle, self.handle, X.ptr, T.ptr, rows,
                                                    C.byref(loss), gp))
        return (loss.value, g) if want_grad else loss.value

    def adam_step(self, lr: float):
        self.ctx.check(self.ctx.L.dt_mlp_adam_step(self.ctx.handle, self.handle, lr))

    def close(self):
        if self.handle:
            self.ctx.L.dt_mlp_destroy(self.ctx.handle, self.handle)
            self.handle = None


def upload_rows(ctx: Context, X: np.ndarray) -> DeviceArray:
    """Upload a host (rows x width) array as a row-major FP32 device buffer."""
    X = np.asarray(X, dtype=np.float32)
    D = ctx.rows_array(X.shape[0], X.shape[1])
    return D.upload(X.T)


def download_rows(D: DeviceArray) -> np.ndarray:
    return D.download().T.copy()


class SpTensor:
    """dt_sp: an N-way COO tensor on the device with its per-mode sorted copies
    (FP32 values, or FP64 values when val is an F64 array)."""

    def __init__(self, ctx: Context, dims, idx, val, nnz: int):
        self.ctx = ctx
        self.dims = tuple(int(d) for d in dims)
        self.nnz = int(nnz)
        h = C.c_void_p()
        nd = len(self.dims)
        dd = (C.c_size_t * nd)(*self.dims)
        ip = (C.c_void_p * nd)(*[a.ptr for a in idx])
        create = ctx.L.dt_sp_create_f64 if val.dtype == DT_F64 else ctx.L.dt_sp_create
        ctx.check(create(ctx.handle, nd, dd, self.nnz, ip, val.ptr, C.byref(h)))
        self.handle = h
        sm, nz, nrm = C.c_int(), C.c_int64(), C.c_double()
        ctx.L.dt_sp_info(h, C.byref(sm), C.byref(nz), C.byref(nrm))
        self.slice_mode, self.norm = sm.value, nrm.value

    @staticmethod
    def from_host(ctx: Context, dims, idx_lists, vals, f64=False):
        idx = [ctx.to_device(np.ascontiguousarray(a, dtype=np.int32)) for a in idx_lists]
        val = ctx.to_device(np.ascontiguousarray(vals, dtype=np.float64 if f64 else np.float32))
        return SpTensor(ctx, dims, idx, val, len(vals))

    def _ut(self, Us):
        """device Ut[m] (R_m x I_m column-major) from host factors U[m] (I_m x R_m)."""
        return [self.ctx.to_device(np.ascontiguousarray(np.asarray(u, dtype=np.float64).T)) for u in Us]

    def ttmc_mtr(self, n: int, Us):
        """Function::ttmc_mTR: L = Y_(n)^T U_n (K x R_n) for host factors Us[m] (I_m x R_m)."""
        ranks = [np.shape(u)[1] for u in Us]
        ut = self._ut(Us)
        K = int(np.prod([r for m, r in enumerate(ranks) if m != n]))
        L = self.ctx.array((K, ranks[n]), DT_F64)
        up = (C.c_void_p * len(Us))(*[u.ptr for u in ut])
        rk = (C.c_size_t * len(Us))(*ranks)
        self.ctx.check(self.ctx.L.dt_sp_ttmc_mtr(self.ctx.handle, self.handle, n, up, rk, L.ptr))
        return L.download()

    def ttmc_mtl(self, n: int, Us, L):
        """Function::ttmc_mTL: (Y_(n) L)^T (R_n x I_n) for host factors and host L (K x R_n)."""
        ranks = [np.shape(u)[1] for u in Us]
        ut = self._ut(Us)
        dL = self.ctx.to_device(np.asarray(L, dtype=np.float64))
        ret = self.ctx.array((ranks[n], self.dims[n]), DT_F64)
        up = (C.c_void_p * len(Us))(*[u.ptr for u in ut])
        rk = (C.c_size_t * len(Us))(*ranks)
        self.ctx.check(self.ctx.L.dt_sp_ttmc_mtl(self.ctx.handle, self.handle, n, up, rk, dL.ptr, ret.ptr))
        return ret.download()

    def hooi_session(self, ranks, seed=0, U_init=None):
        return SpHooiSession(self, ranks, seed, U_init)

    def mode_arrays(self, n: int):
        """(perm, rowptr, nonempty_rows) of mode n, downloaded."""
        pp, rp, nr = C.c_void_p(), C.c_void_p(), C.c_int64()
        self.ctx.check(self.ctx.L.dt_sp_mode_arrays(self.handle, n, C.byref(pp), C.byref(rp), C.byref(nr)))
        perm = np.empty(self.nnz, np.int32)
        rowptr = np.empty(self.dims[n] + 1, np.int32)
        self.ctx.d2h(perm.ctypes.data, pp.value, perm.nbytes)
        self.ctx.d2h(rowptr.ctypes.data, rp.value, rowptr.nbytes)
        return perm, rowptr, nr.value

    def ttmc(self, n: int, Ur):
        """Y_(n) (dims[n] x 256) for row-major FP32 device factors Ur[m] (dims[m] x 16)."""
        Y = self.ctx.rows_array(self.dims[n], 256)
        ptrs = (C.c_void_p * 3)(*[u.ptr if u is not None else None for u in Ur])
        rk = (C.c_size_t * 3)(16, 16, 16)
        self.ctx.check(self.ctx.L.dt_sp_ttmc(self.ctx.handle, self.handle, n, ptrs, rk, Y.ptr))
        return Y

    def hooi(self, max_iter=10, seed=0, want_factors=True, ranks=None, U_init=None):
        nd = len(self.dims)
        if ranks is None:
            ranks = (16,) * nd
        U = [self.ctx.array((d, r), DT_F64) for d, r in zip(self.dims, ranks)] if want_factors else []
        G = self.ctx.array(tuple(ranks), DT_F64) if want_factors else None
        up = (C.c_void_p * nd)(*[u.ptr for u in U]) if U else None
        dinit = [self.ctx.to_device(np.asarray(u, dtype=np.float64)) for u in U_init] if U_init is not None else None
        ip = (C.c_void_p * nd)(*[u.ptr for u in dinit]) if dinit else None
        per = (C.c_double * max_iter)()
        res, ms = C.c_double(), C.c_double()
        rk = (C.c_size_t * nd)(*ranks)
        self.ctx.check(self.ctx.L.dt_hooi_sparse(self.ctx.handle, self.handle, rk, max_iter, seed, ip, up,
                                                  G.ptr if G else None, per, C.byref(res), C.byref(ms)))
        return {"U": [u.download() for u in U], "G": G.download() if G else None,
                "residuals": list(per), "residual": res.value, "sweeps_ms": ms.value}

    def close(self):
        if self.handle:
            self.ctx.L.dt_sp_destroy(self.handle)
            self.handle = None


class SpHooiSession:
    """dt_sp_hooi_create / dt_sp_hooi_sweep / dt_sp_hooi_get: one sweep per call."""

    def __init__(self, sp: SpTensor, ranks, seed=0, U_init=None):
        self.sp, self.ctx, self.ranks = sp, sp.ctx, tuple(int(r) for r in ranks)
        nd = len(sp.dims)
        self._init = [self.ctx.to_device(np.asarray(u, dtype=np.float64)) for u in U_init] if U_init else None
        ip = (C.c_void_p * nd)(*[u.ptr for u in self._init]) if self._init else None
        h = C.c_void_p()
        self.ctx.check(self.ctx.L.dt_sp_hooi_create(self.ctx.handle, sp.handle, (C.c_size_t * nd)(*self.ranks), seed,
                                                    ip, C.byref(h)))
        self.h = h

    def sweep(self):
        """(residual, [||L||_F per mode])"""
        nd = len(self.sp.dims)
        ln = (C.c_double * nd)()
        r = C.c_double()
        self.ctx.check(self.ctx.L.dt_sp_hooi_sweep(self.h, ln, C.byref(r)))
        return r.value, list(ln)

    def result(self):
        U = [self.ctx.array((d, r), DT_F64) for d, r in zip(self.sp.dims, self.ranks)]
        G = self.ctx.array(self.ranks, DT_F64)
        ms = C.c_double()
        self.ctx.check(self.ctx.L.dt_sp_hooi_get(self.h, (C.c_void_p * len(U))(*[u.ptr for u in U]), G.ptr,
                                                 C.byref(ms)))
        return {"U": [u.download() for u in U], "G": G.download(), "sweeps_ms": ms.value}

    def close(self):
        if self.h:
            self.ctx.L.dt_sp_hooi_destroy(self.h)
            self.h = None


def sp_merge_duplicates(ctx: Context, dims, idx_lists, vals):
    """dt_sp_merge_duplicates on host lists: (idx lists, FP64 values) merged."""
    nd = len(dims)
    idx = [ctx.to_device(np.ascontiguousarray(a, dtype=np.int32)) for a in idx_lists]
    val = ctx.to_device(np.ascontiguousarray(vals, dtype=np.float64))
    nu = C.c_int64()
    ctx.check(ctx.L.dt_sp_merge_duplicates(ctx.handle, nd, (C.c_size_t * nd)(*dims), len(vals),
                                           (C.c_void_p * nd)(*[a.ptr for a in idx]), val.ptr, C.byref(nu)))
    k = nu.value
    return [a.download()[:k] for a in idx], val.download()[:k]


def sp_gen_zipf(ctx: Context, dims, ndraws: int, exponent: float, seed: int):
    """(idx DeviceArrays, val DeviceArray, nnz): merged Zipf COO (capacity ndraws)."""
    idx = [ctx.array((ndraws,), DT_I32) for _ in dims]
    val = ctx.array((ndraws,), DT_F32)
    nnz = C.c_int64()
    dd = (C.c_size_t * len(dims))(*dims)
    ip = (C.c_void_p * len(dims))(*[a.ptr for a in idx])
    ctx.check(ctx.L.dt_sp_gen_zipf(ctx.handle, len(dims), dd, ndraws, exponent, seed, ip, val.ptr, C.byref(nnz)))
    return idx, val, nnz.value


def sp_sort_mode(ctx: Context, keys: DeviceArray, nnz: int, dim: int):
    perm = ctx.array((nnz,), DT_I32)
    rowptr = ctx.array((dim + 1,), DT_I32)
    ctx.check(ctx.L.dt_sp_sort_mode(ctx.handle, nnz, keys.ptr, dim, perm.ptr, rowptr.ptr))
    return perm, rowptr
