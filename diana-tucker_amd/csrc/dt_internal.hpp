// Internal declarations of libdiana_hip.so (gfx950).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "diana_hip.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

struct dt_kstat {
    int64_t count = 0;
    double ms = 0.0;
};

struct dt_pending {
    const char *name;
    hipEvent_t a, b;
};

struct dt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool timing = false;
    std::vector<std::string> timing_prefixes;  // when non-empty only launches whose name starts
                                               // with one of these are timed
    std::vector<dt_pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, dt_kstat> stats;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    // Workspace slots: each slot is a device buffer that only grows.
    static constexpr int kSlots = 40;
    void *ws[kSlots] = {};
    size_t ws_size[kSlots] = {};
    // Host-pinned scalar staging area for small device->host results.
    double *pinned = nullptr;
    // resident workgroups on this context's device per kernel (occupancy x CUs)
    std::map<const void *, int64_t> resident_wgs;
    // device address of dt_gram_noalloc's static scratch on this device (set by
    // dt_create through dt_noalloc_setup; null with the reason in noalloc_err)
    void *noalloc_scratch = nullptr;
    std::string noalloc_err;
    // Side work: sums of squares that read only the input tensor, queued by
    // dt_side_sumsq_setup, started on side_stream (default priority) by
    // dt_side_kick where the main stream enters a phase that leaves HBM idle,
    // and waited for by dt_side_join before their results are read.
    hipStream_t side_stream = nullptr;
    hipEvent_t side_fork = nullptr, side_done = nullptr;
    struct {
        int count = 0;
        bool kicked = false;
        const void *src[2] = {};
        int dt[2] = {};
        size_t len[2] = {};
        double *out_dev = nullptr;  // count doubles
    } side;
};
// looks up this context's device address of the allocation-free scratch
int dt_noalloc_setup(dt_ctx *ctx);

// Workspace slot ids (one owner each, so nested helpers never alias).
enum dt_ws_slot {
    WS_TTM_M = 0,      // converted/transposed factor for dt_ttm
    WS_GRAM_PART,      // gram split-K partials
    WS_SUMSQ,          // sumsq block partials
    WS_GEMM_PART,      // gemm split-K partials
    WS_SMALL0,         // small FP64 scratch (R x R)
    WS_SMALL1,
    WS_SMALL2,
    WS_SMALL3,
    WS_QR_T0,          // QR working matrix
    WS_QR_T1,
    WS_EIG_X,          // subspace iteration
    WS_EIG_Y,
    WS_EIG_H,
    WS_STATUS,         // device int status word(s)
    WS_TTT,            // ttt_except staging
    WS_GEN,            // generator scratch
    WS_GEN2,
    WS_TTMC0,          // dt_ttmc intermediates
    WS_TTMC1,
    WS_TTMC_M,         // dt_ttmc fused-pair factor views
    WS_SPARSE0,
    WS_SPARSE1,
    WS_SPARSE2,
    WS_SPARSE3,
    WS_MLP0,
    WS_MLP1,
    WS_MLP2,
    WS_SKETCH,         // learned start: every mode's sketch, Gram and eigenvectors
    WS_CHOLQR,         // dt_qr_f64's multi-workgroup Cholesky-QR: partial Grams, R^{-1}, flag
    WS_TTM_SWZ,        // mode-0 BF16 product: the factor in MFMA operand order
    WS_SKPRE,          // learned start: the Kronecker sketch prefix modes 0 and 1 share
    WS_SIDE0,          // side work: block partials of the first and second sum, results
    WS_SIDE1,
    WS_SIDE_OUT,
    WS_COUNT_
};
static_assert(WS_COUNT_ <= dt_ctx::kSlots, "too many workspace slots");

int dt_fail(dt_ctx *ctx, int code, const char *fmt, ...);
// workgroups of `threads` threads of `kernel` resident on the context's device
// at once (cached per context); 0 when the query fails
template <typename K>
int64_t dt_resident_wgs(dt_ctx *ctx, K kernel, int threads) {
    const void *key = (const void *)kernel;
    auto it = ctx->resident_wgs.find(key);
    if (it != ctx->resident_wgs.end()) return it->second;
    int cus = 0, blocks = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, threads, 0) != hipSuccess) return 0;
    const int64_t n = (int64_t)cus * (blocks < 1 ? 1 : blocks);
    ctx->resident_wgs[key] = n;
    return n;
}
void *dt_ws(dt_ctx *ctx, int slot, size_t bytes);  // nullptr on failure (error set)
hipEvent_t dt_get_event(dt_ctx *ctx);
void dt_collect_timing(dt_ctx *ctx);

#define DT_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return dt_fail(ctx, DT_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,       \
                           hipGetErrorString(e_));                                        \
    } while (0)

#define DT_TRY(call)                   \
    do {                               \
        int s_ = (call);               \
        if (s_ != DT_OK) return s_;    \
    } while (0)

#define DT_CHECK_ARG(cond, ...)                                  \
    do {                                                         \
        if (!(cond)) return dt_fail(ctx, DT_EINVAL, __VA_ARGS__); \
    } while (0)

// Brackets one kernel launch with events when timing is on.
struct dt_ktimer {
    dt_ctx *c;
    const char *name;
    hipEvent_t a = nullptr;
    bool on = false;
    dt_ktimer(dt_ctx *c_, const char *n) : c(c_), name(n) {
        on = c->timing && c->timing_prefixes.empty();
        if (c->timing && !on)
            for (const std::string &p : c->timing_prefixes)
                if (strncmp(n, p.c_str(), p.size()) == 0) on = true;
        if (on) {
            a = dt_get_event(c);
            hipEventRecord(a, c->stream);
        }
    }
    ~dt_ktimer() {
        if (on) {
            hipEvent_t b = dt_get_event(c);
            hipEventRecord(b, c->stream);
            c->pending.push_back({name, a, b});
        }
    }
};

#define DT_LAUNCH(name, kernel, grid, block, shmem, ...)                              \
    do {                                                                              \
        dt_ktimer kt_(ctx, name);                                                     \
        hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, __VA_ARGS__);    \
    } while (0);                                                                      \
    DT_HIP(hipGetLastError())

// ---- bf16 helpers (storage = uint16_t, round to nearest even) ----------------
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    return __builtin_bit_cast(float, ((uint32_t)h) << 16);
}
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN stays NaN
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

inline size_t dt_elem_size(int dt) {
    switch (dt) {
        case DT_F64: return 8;
        case DT_F32: return 4;
        case DT_BF16: return 2;
        case DT_I32: return 4;
        default: return 0;
    }
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- internal entry points shared between translation units -----------------
// Generic mode-n product with a factor given in both layouts:
//   Ukj : K x J column-major (k fastest)   -- "U layout"
//   Mjk : J x K column-major (j fastest)   -- "M layout" (the reference's M)
// For F32/BF16 inputs both are FP32 (BF16 A uses Ub16, K x J bf16).
struct dt_factor_views {
    const float *U32;     // K x J
    const float *M32;     // J x K
    const uint16_t *Ub16; // K x J (only for BF16 A)
    const double *M64;    // J x K (only for F64 A)
};
// stream = true: A is read exactly once (the full input tensor); its loads are
// non-temporal and the launch is timed under the "<kernel>_stream" name.
int dt_ttm_views(dt_ctx *ctx, const void *A, int dtA, int64_t F, int64_t K, int64_t B,
                 const dt_factor_views &fv, int64_t J, void *Y, int dtY, bool stream = false);
int dt_gram_internal(dt_ctx *ctx, const void *A, int dtA, int64_t F, int64_t I, int64_t B,
                     double *S);
int dt_sumsq_dev(dt_ctx *ctx, const void *A, int dt, size_t n, double *out_host);
// Side work (dt_ctx::side): queue the sums of squares of up to two tensors;
// start them (no-op when none is queued or they already run); wait for them
// and copy the sums to out_host (nullptr discards them).
int dt_side_sumsq_setup(dt_ctx *ctx, int count, const void *const *src, const int *dt, const size_t *len);
int dt_side_kick(dt_ctx *ctx);
int dt_side_join(dt_ctx *ctx, double *out_host);
int dt_transpose_f64(dt_ctx *ctx, double *B, const double *A, size_t m, size_t n);
int dt_f64_to_views(dt_ctx *ctx, const double *U64, size_t K, size_t J, float *U32, float *M32,
                    uint16_t *Ub16, double *M64);
int dt_chol_upper(dt_ctx *ctx, const double *G, size_t n, double *R, double *Rinv, int *info_host);
int dt_gemm_f64_internal(dt_ctx *ctx, int tA, int tB, size_t m, size_t n, size_t k, double alpha,
                         const double *A, size_t lda, const double *B, size_t ldb, double beta,
                         double *C, size_t ldc);
int dt_jacobi_eig(dt_ctx *ctx, double *H, size_t n, double *V, double *w);
int dt_jacobi_eig_batched(dt_ctx *ctx, double *H, size_t n, int count, double *V, double *w);
// C (IA x JB, FP64) = A_(n) B_(n)ᵀ, FP32 or bf16 tensors of mode-n sizes IA and JB
int dt_cross_internal(dt_ctx *ctx, const void *A, const void *B, int dt, int64_t F, int64_t IA, int64_t JB,
                      int64_t Bk, double *C);
int dt_copy_f64(dt_ctx *ctx, double *dst, const double *src, size_t n);
int dt_fill_normal_f64(dt_ctx *ctx, double *out, size_t n, uint64_t seed, uint64_t stream);
// fused single-workgroup FP64 steps (dt_small.hip); failures OR bits into *flag
int dt_als_prep(dt_ctx *ctx, const double *L, int I, int R, double *P, double *work, int *flag);
int dt_als_finish(dt_ctx *ctx, const double *P, const double *W, int I, int R, double *L,
                  double *work, int *flag);
int dt_cholqr3_dev(dt_ctx *ctx, double *X, int m, int n, double *work, int *flag);
// up to DT_MAX_NDIM independent (XᵀX)^{-1} in one launch (one workgroup each), n <= 64
struct dt_gram_inverse_batch {
    const void *X[DT_MAX_NDIM];
    int dt[DT_MAX_NDIM], m[DT_MAX_NDIM], n[DT_MAX_NDIM];
    double *C[DT_MAX_NDIM];
    int count;
    int ch;  // rows per staged chunk (set by dt_gram_inverse_batch_dev)
};
int dt_gram_inverse_batch_dev(dt_ctx *ctx, const dt_gram_inverse_batch &b, int *flag);
// C = G^{-1}; above order 64 work must hold n ints and C must not alias G
int dt_inverse_flag(dt_ctx *ctx, const double *G, int n, double *C, double *work, int *flag, int bit);
int dt_orth_dev(dt_ctx *ctx, double *X, size_t m, size_t n, double *work, int *flag);
// Rinv (n x n upper, column-major) = inverse Cholesky factor of G (+ the
// shifted-CholeskyQR shift for m_rows rows when shifted); failure ORs 4 into *flag
int dt_chol_rinv_dev(dt_ctx *ctx, const double *G, int n, bool shifted, double m_rows, double *Rinv, int *flag);
// The same for any order: above 64 the factorisation runs in place in G
// (global, overwritten) with invd (n doubles) as scratch
int dt_chol_rinv_ws(dt_ctx *ctx, double *G, int n, bool shifted, double m_rows, double *Rinv, double *invd,
                    int *flag);
// out[e] = sum over p < nparts of part[p * ne + e] in a fixed order
int dt_part_reduce(dt_ctx *ctx, const double *part, int nparts, int ne, double *out);
size_t dt_als_ws_doubles(size_t I, size_t R);
// X (I x R, R <= 64) <- its orthonormal QR factor (positive R diagonal) by
// shifted Cholesky-QR3 over many workgroups; no host wait, breakdown (and, with
// rank_check, a rank-deficient X) ORs 4 into *flag;
// ws: (ceil(I / 128) + 2) R^2 + R + 1 doubles
int dt_cholqr3_rows(dt_ctx *ctx, double *X, int I, int R, double *ws, int *flag, bool rank_check = false);
// doubles of the workspace dt_cholqr3_rows takes for an I x R panel
size_t dt_cholqr3_ws_doubles(size_t I, size_t R);
int dt_als_update(dt_ctx *ctx, const double *S, double *U, int I, int R, int inner, double *ws,
                  int *flag, float *U32, float *M32, uint16_t *Ub16);
int dt_als_fused(dt_ctx *ctx, const double *S, double *U, int I, int R, int inner, double *L,
                 double *P, double *W, int *flag);
// T (J0 x J1 x B) = A (K0 x K1 x B) x_0 U0ᵀ x_1 U1ᵀ in one pass over A.
// U0: K0 x J0 (FP32, or bf16 for a BF16 A); M1: J1 x K1 FP32.  Returns
// DT_EUNSUPPORTED when the shape does not fit the fused kernel.
int dt_ttm2_lead(dt_ctx *ctx, const void *A, int dtA, int64_t K0, int64_t K1, int64_t B, const void *U0,
                 const float *M1, int64_t J0, int64_t J1, float *T);
// T (F x J2 x J3) = A (F x K2 x K3) x_1 M2 x_2 M3 (FP32, M = J x K): the last two
// modes in one pass; DT_EUNSUPPORTED when the shape does not fit
int dt_ttm2_trail(dt_ctx *ctx, const float *A, int dtA, int64_t F, int64_t K2, int64_t K3, const float *M2,
                  const float *M3, int64_t J2, int64_t J3, float *T);
int dt_ttm_f64acc(dt_ctx *ctx, const void *A, int dtA, int64_t F, int64_t K, int64_t B,
                  const double *M64, int64_t J, double *Y);
int dt_dot_f64_dev(dt_ctx *ctx, const double *a, const double *b, size_t n, double *out_dev);
// ---- sort and scan (dt_sort.hip) ----------------------------------------------
// out[i] = sum of in[0..i) (int32; in and out may alias); tmp: dt_scan_tmp_bytes(n)
int dt_exclusive_scan_i32(dt_ctx *ctx, const int32_t *in, int32_t *out, int64_t n, int32_t *tmp);
size_t dt_scan_tmp_bytes(int64_t n);
// stable LSD radix sort of (key, payload) by the low nbits of the key, 8 bits a
// pass; kin != kout and vin != vout; tmp: dt_radix_sort_tmp_bytes(n, sizeof key)
size_t dt_radix_sort_tmp_bytes(int64_t n, int key_bytes);
int dt_radix_sort_u32(dt_ctx *ctx, const uint32_t *kin, uint32_t *kout, const int32_t *vin, int32_t *vout, int64_t n,
                      int nbits, void *tmp);
int dt_radix_sort_u64(dt_ctx *ctx, const uint64_t *kin, uint64_t *kout, const int32_t *vin, int32_t *vout, int64_t n,
                      int nbits, void *tmp);
