/*
 * diana_hip.h -- C ABI of the MI355X (gfx950) HOOI-ALS library `libdiana_hip.so`.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and returns a
 * status code (DT_OK = 0).  No entry point calls exit()/abort(); the text of the
 * last failure is available from dt_last_error().  Tensors are column-major with
 * the first index fastest, exactly as the reference stores them
 * (template/tensor_mpi.tpp:49-53).  Matrices are column-major with implicit
 * leading dimension unless an ld argument is given.
 *
 * Each entry point names the reference interface it replaces.  The reference
 * calls these through `A.op()->...` (an Operator<Ty> instance owned by each
 * Tensor, include/operator.hpp:8-70) and through the Function:: templates
 * (include/function.hpp:6-94); the C++ mirror in diana-tucker_amd/host keeps
 * those names and forwards here.
 *
 * Device memory: pointers marked "device" are allocations made with dt_malloc()
 * (or any hipMalloc'd memory on the context's device).  All work is enqueued on
 * the context's HIP stream; entry points that return host values synchronise
 * that stream.  dt_hooi_create also runs the tensor norms on a second stream of
 * the context (default priority) beside the start's eigen step and waits for
 * them before it returns, so no work of a call is left on that stream.  A
 * context must be used from one host thread at a time (the reference is driven
 * by one thread, spmain.cpp:10-15).
 */
#ifndef DIANA_HIP_H
#define DIANA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    DT_OK = 0,
    DT_EINVAL = 1,       /* bad argument (shape, rank > mode size, null pointer) */
    DT_ENOMEM = 2,       /* device allocation failed */
    DT_EHIP = 3,         /* HIP runtime error */
    DT_ENUMERIC = 4,     /* non-finite value or breakdown (singular Gram, ||A|| = 0) */
    DT_EUNSUPPORTED = 5  /* combination not implemented */
};

/* ---- element types ----------------------------------------------------- */
enum { DT_F64 = 0, DT_F32 = 1, DT_BF16 = 2, DT_I32 = 3 };

#define DT_MAX_NDIM 8

typedef struct dt_ctx dt_ctx;

/* ---- context, memory, stream ------------------------------------------ */
/* Creates a context on HIP device `device` with its stream (and the side
 * stream dt_hooi_create uses internally).  Replaces the
 * per-tensor `new Operator<Ty>()` (template/tensor_mpi.tpp:26) and the
 * communicator set-up (src/communicator.cpp:3-15) of a single rank. */
int dt_create(dt_ctx **ctx, int device);
int dt_destroy(dt_ctx *ctx);
const char *dt_last_error(const dt_ctx *ctx);
/* Operator<Ty>::alloc / free (template/operator/operator_cpu.tpp:6-12). */
int dt_malloc(dt_ctx *ctx, void **ptr, size_t bytes);
int dt_free(dt_ctx *ctx, void *ptr);
/* Operator<Ty>::mcpy (operator_cpu.tpp:14-17), split by direction. */
int dt_h2d(dt_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
int dt_d2h(dt_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
int dt_d2d(dt_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes);
/* Operator<Ty>::constant(A, 0, n) (src/operator/operator_double.cpp:51-56). */
int dt_zero(dt_ctx *ctx, void *dev, size_t bytes);
int dt_sync(dt_ctx *ctx);
/* Element type conversion dst[i] = (dt_dst) src[i] on the device (F64/F32/BF16). */
int dt_convert(dt_ctx *ctx, void *dst, int dt_dst, const void *src, int dt_src, size_t n);

/* ---- timing ------------------------------------------------------------
 * With timing enabled every kernel launch of the library is bracketed by HIP
 * events on the context stream; dt_timing_get returns the launch count and the
 * summed event time for one kernel name (dt_timing_list lists the names seen).
 * dt_timer_* time a region of the stream with a pair of events.  Replaces the
 * MPI_Wtime-based Summary timers (src/summary.cpp:29-69). */
int dt_timing_enable(dt_ctx *ctx, int on);
int dt_timing_reset(dt_ctx *ctx);
/* Only launches whose kernel name starts with one of the comma-separated
 * prefixes are bracketed (NULL or "" = all); keeps the event overhead off
 * the kernels that are not measured. */
int dt_timing_filter(dt_ctx *ctx, const char *prefix);
int dt_timing_get(dt_ctx *ctx, const char *kernel, int64_t *count, double *total_ms);
int dt_timing_list(dt_ctx *ctx, char *buf, size_t buflen); /* "name count ms\n" lines */
int dt_timer_start(dt_ctx *ctx);
int dt_timer_stop(dt_ctx *ctx, double *elapsed_ms);

/* ---- operator level ----------------------------------------------------- */
/* Y = A x_n M.  Replaces Function::ttm (template/function/tensor.tpp:272-335)
 * = Operator::tenmatt + matmulNT + mattten (operator_cpu.tpp:83-135,
 * operator_double.cpp:176-194) at one rank, without the unfold copies.
 * A: device, `ndim` modes of sizes shape[], element type dtA (F64/F32/BF16).
 * M: device, J x shape[n] column-major, element type dtM (F64/F32).
 * Y: device, shape with shape[n] replaced by J, element type dtY (F64/F32).
 * F64 A requires F64 Y and runs in FP64; F32/BF16 A accumulate in FP32. */
int dt_ttm(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape,
           const void *M, int dtM, size_t J, int n, void *Y, int dtY);

/* Y = A x_{idx[0]} M[0] x_{idx[1]} M[1] ... x_{idx[nm-1]} M[nm-1].
 * Replaces Function::ttmc (template/function/tensor.tpp:337-348), which
 * applies Function::ttm in sequence.  M[s]: device, J[s] x (current size of
 * mode idx[s]) column-major, all of type dtM (F64/F32).  dtY is F64 for an
 * F64 tensor, F32 otherwise.  A chain that starts with modes 0 and 1 of an
 * F32/BF16 tensor runs those two products in one pass over A (no
 * intermediate of A's size is written) when shape[0] % 64 (F32) or % 128
 * (BF16), shape[1] % 64 == 0, J <= 32 for mode 0 and J <= 64 for mode 1. */
int dt_ttmc(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape,
            const void *const *M, int dtM, const size_t *J, const int *idx, int nm,
            void *Y, int dtY);

/* S = A_(n) A_(n)^T (shape[n] x shape[n], column-major, dtS F32 or F64).
 * Replaces Function::gram(A, n) (template/function/tensor.tpp:18-120) =
 * tenmat + matmulNT at one rank.  FP32/BF16 inputs accumulate in FP32 on
 * MFMA; F64 input accumulates in FP64. */
int dt_gram(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape, int n,
            void *S, int dtS);

/* dt_gram for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  No
 * counterpart in the reference.  Same arguments and results as dt_gram; the
 * split-K partials (and, for an F32 S, the FP64 Gram before its conversion)
 * live in a static device array of DT_GRAM_NOALLOC_SCRATCH bytes, one per GPU,
 * held from library load to process exit, instead of the context workspace.
 * Calls no hipMalloc/hipFree and creates no events (its launches are not
 * timed).  A Gram whose scratch exceeds the maximum returns DT_EINVAL: for
 * FP32/BF16 input an order above 2048 (F64 S) or 1182 (F32 S); for FP64 input
 * an order above 1448 with an F32 S (an F64 S needs no scratch).  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (two streams, two contexts, two threads) race. */
#define DT_GRAM_NOALLOC_SCRATCH ((size_t)16 << 20)
int dt_gram_noalloc(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape, int n,
                    void *S, int dtS);

/* C = A_(n) B_(n)^T where A and B agree on every mode except n (C is
 * shapeA[n] x shapeB[n], FP64).  Replaces Function::ttt_except
 * (template/function/tensor.tpp:132-259).  A and B are F64 or F32. */
int dt_ttt_except(dt_ctx *ctx, const void *A, const void *B, int dt, int ndim,
                  const size_t *shapeA, size_t JB, int n, double *C);

/* sum of squares in FP64 of n elements (host result).  sqrt of it replaces
 * Operator<double>::fnorm (src/operator/operator_double.cpp:74-82). */
int dt_sumsq(dt_ctx *ctx, const void *A, int dt, size_t n, double *out_host);

/* ---- FP64 small dense (device pointers) --------------------------------- */
/* C = alpha op(A) op(B) + beta C, column-major with leading dimensions.
 * Replaces Operator<double>::matmulNN/NT/TN (operator_double.cpp:157-216),
 * which call cblas_dgemm with implicit leading dimensions. */
int dt_gemm_f64(dt_ctx *ctx, int transA, int transB, size_t m, size_t n, size_t k,
                double alpha, const double *A, size_t lda, const double *B, size_t ldb,
                double beta, double *C, size_t ldc);
/* Reduced QR of an m x n (m >= n) matrix: Q m x n with orthonormal columns,
 * R n x n upper triangular with positive diagonal (R may be NULL).
 * Replaces Operator<double>::QR (operator_double.cpp:131-155, dgeqrf+dorgqr);
 * the column signs differ from LAPACK's, the spans are equal. */
int dt_qr_f64(dt_ctx *ctx, double *Q, double *R, const double *A, size_t m, size_t n);
/* C = A^{-1}, n x n.  Replaces Operator<double>::inverse (operator_double.cpp:84-103). */
int dt_inverse_f64(dt_ctx *ctx, double *C, const double *A, size_t n);
/* Leading r eigenpairs of the symmetric n x n matrix S (descending): V n x r,
 * w r values (device).  No reference counterpart; used for HOSVD start factors. */
int dt_syev_top_f64(dt_ctx *ctx, const double *S, size_t n, size_t r, double *V, double *w);
/* Row orthonormalisation of a len1 x len2 column-major matrix (rows made
 * orthonormal, positive diagonal of the triangular factor).  Replaces
 * Operator<double>::orth2 (operator_double.cpp:248-277); orth1 (:218-246) is
 * dt_qr_f64 with R = NULL. */
int dt_orth_rows_f64(dt_ctx *ctx, double *A, size_t len1, size_t len2);

/* ---- seeded generators --------------------------------------------------
 * Counter-based Philox4x32-10; element i of stream `stream` is a function of
 * (seed, stream, i) only, so host and device generate identical values. */
int dt_gen_normal(dt_ctx *ctx, void *out, int dt, size_t n, uint64_t seed, uint64_t stream);
int dt_gen_uniform(dt_ctx *ctx, void *out, int dt, size_t n, uint64_t seed, uint64_t stream);

/* Synthetic dense Tucker model A = G x_1 U_1 ... x_N U_N + E with
 * ||E|| = noise * ||X||.  G ~ N(0,1) (then scaled by decay^k along every
 * mode when decay > 0), U_n = Q of the QR of an N(0,1) I_n x R_n matrix.
 * Writes A in F32 (A_f32) and/or BF16 (A_bf16); either may be NULL.  The true
 * factors (FP64, I_n x R_n) are written to U_true[n] when non-NULL. */
typedef struct dt_tucker_spec {
    int32_t ndim;
    size_t shape[DT_MAX_NDIM];
    size_t ranks[DT_MAX_NDIM];
    double noise;     /* relative Frobenius norm of E */
    double decay;     /* geometric decay of the core along each mode (0 = none) */
    uint64_t seed;
} dt_tucker_spec;
int dt_gen_tucker(dt_ctx *ctx, const dt_tucker_spec *spec, float *A_f32, uint16_t *A_bf16,
                  double *const *U_true, double *normX_host, double *normA_host);

/* ---- sparse ---------------------------------------------------------------
 * Sparse HOOI of Algorithm::SpTucker::Sp_HOOI_ALS
 * (template/algorithm/sp_tucker/sp_hooi_als.tpp:67-118) for any number of
 * modes: random orthonormal start, then per sweep and mode n one step
 * U_n <- orth(Y_(n) Y_(n)ᵀ U_n) (Sp_TTMC_ALS_, :48-65 = ttmc_mTR + ttmc_mTL +
 * orth2), core = Y_(N-1)ᵀ U_{N-1} (:110).  Indices int32, values FP32 (or FP64
 * with dt_sp_create_f64), all device pointers.  Y_(n) columns follow the
 * reference's Kronecker order (add_outer_product_all, sp_tensor.tpp:367-417):
 * the other modes ascending, the first fastest.  3-way tensors with FP32
 * values and every rank <= 16 take the MFMA TTMc pipeline; other shapes the
 * general kernels (every rank <= 64, product of the other modes' ranks <= 4096),
 * in FP64 when the tensor holds FP64 values.  Every rank must be at most its
 * mode's size and at most the product of the other ranks (Y_(n) has that many
 * columns); else DT_EINVAL. */
typedef struct dt_sp dt_sp;
/* COO from independent per-mode truncated Zipf draws P(k) ~ (k+1)^-s (inverse
 * CDF on an FP64 table; Philox uniform streams 7000 + n, values stream 7100,
 * U(0,1] in FP32); duplicates merged: entries in increasing composite-key
 * order (mode 0 in the lowest bits), value = FP64 sum of the draws in draw
 * order rounded to FP32.  idx_out[n] and val_out hold ndraws entries;
 * *nnz_out receives the merged count.  No reference counterpart (the
 * reference reads .tns files, src/read_data.cpp:36-129). */
int dt_sp_gen_zipf(dt_ctx *ctx, int ndim, const size_t *dims, int64_t ndraws, double exponent,
                   uint64_t seed, int32_t *const *idx_out, float *val_out, int64_t *nnz_out);
/* Stable sort of nonzeros by one index list: perm (nnz) and row pointers
 * (dim + 1).  The SpTensor constructor's counting sort (sp_tensor.hpp:117-147)
 * as a stable LSD radix sort: perm is identical. */
int dt_sp_sort_mode(dt_ctx *ctx, int64_t nnz, const int32_t *keys, size_t dim, int32_t *perm,
                    int32_t *rowptr);
/* SpTensor(data_buffer*) (sp_tensor.hpp:72-240): copies the COO and builds a
 * stably sorted copy per mode; slice mode = the largest mode, first on ties. */
int dt_sp_create(dt_ctx *ctx, int ndim, const size_t *dims, int64_t nnz, const int32_t *const *idx,
                 const float *val, dt_sp **sp);
/* SpTensor<double>(data_buffer*): the same with FP64 values, kept for the FP64
 * path (TTMc, L and the factor update accumulate in FP64); an FP32 rounding
 * of them serves nothing else. */
int dt_sp_create_f64(dt_ctx *ctx, int ndim, const size_t *dims, int64_t nnz, const int32_t *const *idx,
                     const double *val, dt_sp **sp);
/* Ingest step after read_data (src/read_data.cpp:80-127, which keeps every
 * line): nonzeros with equal coordinates are merged in place, values summed
 * in FP64 in input order at the first occurrence; survivors keep their input
 * order (a list without duplicates is unchanged).  *nnz_out receives the new
 * count; indices outside [0, dims) give DT_EINVAL. */
int dt_sp_merge_duplicates(dt_ctx *ctx, int ndim, const size_t *dims, int64_t nnz, int32_t *const *idx,
                           double *val, int64_t *nnz_out);
int dt_sp_destroy(dt_sp *sp);
int dt_sp_info(const dt_sp *sp, int *slice_mode, int64_t *nnz, double *norm);
int dt_sp_mode_arrays(const dt_sp *sp, int n, const int32_t **perm, const int32_t **rowptr,
                      int64_t *nonempty_rows);
/* Y_(n) of A x_{m != n} U_mᵀ: Y (shape[n] x 256, row-major FP32; row i =
 * sum over the nonzeros of slice i of val * U_a[i_a,:] ⊗ U_b[i_b,:], entry
 * r_a + 16 r_b, a < b the other modes), Ur[m] row-major FP32 shape[m] x 16.
 * Replaces the Kronecker accumulation of ttmc_mTR / ttmc_mTL
 * (sp_tensor.tpp:236-365, add_outer_product_all :367-417); ranks 16 only. */
int dt_sp_ttmc(dt_ctx *ctx, dt_sp *sp, int n, const float *const *Ur, const size_t *ranks, float *Y);
/* Function::ttmc_mTR (include/function.hpp:70, template/function/
 * sp_tensor.tpp:236-286) as intended: L = Y_(n)ᵀ U_n, K x R_n column-major FP64
 * with K = prod_{m != n} R_m in the Kronecker order above.  Ut[m]: R_m x I_m
 * column-major FP64 (the reference's transposed factors), every mode's. */
int dt_sp_ttmc_mtr(dt_ctx *ctx, dt_sp *sp, int n, const double *const *Ut, const size_t *ranks, double *L);
/* Function::ttmc_mTL (include/function.hpp:74, sp_tensor.tpp:288-365) as
 * intended: ret (R_n x I_n column-major FP64) = (Y_(n) L)ᵀ for L (K x R_n
 * column-major); slices without nonzeros give zero columns. */
int dt_sp_ttmc_mtl(dt_ctx *ctx, dt_sp *sp, int n, const double *const *Ut, const size_t *ranks, const double *L,
                   double *ret);
/* The Sp_HOOI_ALS loop one sweep at a time (sp_hooi_als.tpp:98-109), for the
 * host layer's per-sweep lines: create takes the start as dt_hooi_sparse does;
 * a sweep returns each mode's ||L||_F (what Sp_TTMC_ALS_ prints as ||G||_F,
 * :51-53; ndim entries, or NULL) and the residual sqrt(1 - ||G||^2/||A||^2) of
 * the core after the sweep; get copies the factors (I_n x R_n column-major
 * FP64), the core (R_0 x ... x R_{N-1} column-major FP64) and the sweeps'
 * device time. */
typedef struct dt_sp_hooi dt_sp_hooi;
int dt_sp_hooi_create(dt_ctx *ctx, dt_sp *sp, const size_t *ranks, uint64_t seed, const double *const *U_init,
                      dt_sp_hooi **h);
int dt_sp_hooi_sweep(dt_sp_hooi *h, double *L_norms, double *residual);
int dt_sp_hooi_get(dt_sp_hooi *h, double *const *U_out, double *G_out, double *sweeps_ms);
int dt_sp_hooi_destroy(dt_sp_hooi *h);
/* Sparse HOOI: ranks as above, max_iter sweeps.  Start: U_init[n] (device
 * FP64 shape[n] x R_n column-major, orthonormalised here), or, when U_init is
 * NULL, Philox normals (streams 200 + n).  U_out[n] (shape[n] x R_n FP64
 * column-major) and G_out (R_0 x ... x R_{N-1} FP64 column-major) may be NULL;
 * residual_per_sweep holds max_iter entries (or NULL); sweeps_ms = device
 * time of the sweeps. */
int dt_hooi_sparse(dt_ctx *ctx, dt_sp *sp, const size_t *ranks, int max_iter, uint64_t seed,
                   const double *const *U_init, double *const *U_out, double *G_out,
                   double *residual_per_sweep, double *residual, double *sweeps_ms);

/* ---- learned start ------------------------------------------------------
 * No counterpart in the reference (it starts from random factors,
 * hooi_als.tpp:76-81): the learned initialisation that replaces the HOSVD
 * start.  Per mode n, a Gaussian sketch Y = A_(n) Ω (Ω: prod_{m!=n} I_m x S,
 * Philox stream 5000 + n) gives features X = sqrt(I_n) / σ_1 · Y V (rows of the
 * sketch in the frame of its right singular vectors V, ordered by singular
 * value σ); a row-wise MLP maps each row of X (S values) to R values,
 *     z = x[:R] + W3ᵀ relu(W2ᵀ relu(W1ᵀ x + b1) + b2) + b3,
 * and the start factor is Q of QR(Z).  Weights FP32 (training master copy),
 * rounded to bf16 in every product; products on bf16 MFMA with FP32
 * accumulation.  Parameter order: W1 (S x H, row-major), b1, W2 (H x H), b2,
 * W3 (H x R), b3. */
typedef struct dt_mlp dt_mlp;
int dt_mlp_create(dt_ctx *ctx, dt_mlp **mlp, int d_in, int hidden, int d_out, uint64_t seed);
int dt_mlp_destroy(dt_ctx *ctx, dt_mlp *mlp);
int dt_mlp_shape(const dt_mlp *mlp, int *d_in, int *hidden, int *d_out, int64_t *nparam);
int dt_mlp_get_params(dt_ctx *ctx, const dt_mlp *mlp, float *host, int64_t n);
int dt_mlp_set_params(dt_ctx *ctx, dt_mlp *mlp, const float *host, int64_t n);
/* Z (rows x d_out) = MLP(X (rows x d_in)); device, row-major FP32. */
int dt_mlp_forward(dt_ctx *ctx, dt_mlp *mlp, const float *X, int64_t rows, float *Z);
/* loss = sum_rows |MLP(x) - t|^2 / rows and its gradient (kept for
 * dt_mlp_adam_step; copied to grad_host, nparam floats, when non-NULL). */
int dt_mlp_loss_grad(dt_ctx *ctx, dt_mlp *mlp, const float *X, const float *T, int64_t rows,
                     double *loss, float *grad_host);
int dt_mlp_adam_step(dt_ctx *ctx, dt_mlp *mlp, float lr);
/* Sketch features of mode n: X (shape[n] x S, device, row-major FP32). */
int dt_sketch_features(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape, int n,
                       int S, uint64_t seed, float *X);
/* The same features from the Kronecker sketch Ω = ⊗_{m != n} Ω_m (Ω_m
 * Gaussian shape[m] x q, Philox stream 6000 + m), S = q^(ndim-1): one chain
 * of mode products (dt_ttmc) instead of a dense (prod_{m!=n} I_m) x S Ω. */
int dt_sketch_features_kron(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape,
                            int n, int q, uint64_t seed, float *X);
enum { DT_SKETCH_GAUSSIAN = 0, DT_SKETCH_KRONECKER = 1 };
/* Learned start factors U[n] (device FP64, shape[n] x ranks[n], orthonormal);
 * mlp d_in = sketch width S (= q^(ndim-1) for the Kronecker sketch),
 * d_out = every ranks[n]. */
int dt_learned_start(dt_ctx *ctx, dt_mlp *mlp, const void *A, int dtA, int ndim,
                     const size_t *shape, const size_t *ranks, int sketch, uint64_t seed,
                     double *const *U);

/* ---- dense HOOI --------------------------------------------------------- */
enum {
    DT_INIT_GIVEN = 0,   /* U_init[n] are orthonormal start factors */
    DT_INIT_QR = 1,      /* U_init[n] are raw matrices, orthonormalised first
                            (the reference's randn + reduced_QR, hooi_als.tpp:76-81) */
    DT_INIT_HOSVD = 2,   /* leading R_n eigenvectors of A_(n) A_(n)^T */
    DT_INIT_RANDOM = 3,  /* Philox N(0,1) matrices from `seed`, then QR */
    DT_INIT_LEARNED = 4  /* sketch -> MLP -> QR; falls back to HOSVD (see dt_hooi_result.init_used) */
};
enum {
    DT_SCHED_REFERENCE = 0, /* per mode a full ascending TTM chain, core by a
                               further chain: N+1 full-tensor passes per sweep */
    DT_SCHED_TREE = 1       /* two-level dimension tree: 2 full-tensor passes per
                               sweep, core from the last Y */
};
enum {
    DT_FLAG_DIRECT_RESIDUAL = 1, /* after the last sweep compute ||A - [[G;U]]|| in FP64 */
    DT_FLAG_FP32_POLISH = 2,     /* mixed precision: finish with one FP32 sweep on A_f32 */
    DT_FLAG_REFERENCE_ALS = 8,   /* per-mode update exactly as ALS_ writes it (LU inverses
                                    of LᵀL and PᵀSP); default: the span-equivalent
                                    Cholesky-normalised iteration */
    DT_FLAG_FP64_FIT = 4,        /* form the core for the residual with FP64 accumulation
                                    (one extra pass over A per sweep); needed when the
                                    residual is below ~1e-3, where FP32 rounding of the
                                    core dominates 1 - ||G||^2/||A||^2 */
    DT_FLAG_LEARNED_CHECK = 16,  /* DT_INIT_LEARNED: also run the first sweep from the HOSVD
                                    start and keep whichever fits better (the fallback) */
    DT_FLAG_SKETCH_KRON = 32,    /* DT_INIT_LEARNED: Kronecker sketch (model input q^(N-1)) */
    DT_FLAG_LEARNED_ACCEPT = 64  /* DT_INIT_LEARNED: accept the learned start only when the sweep
                                    after its first one moves the fit by less than
                                    DT_LEARNED_ACCEPT_FIT (HOOI from it has converged to the fit
                                    tolerance), else fall back to the HOSVD start; costs one sweep
                                    (none when max_iter >= 2), not a HOSVD */
};
#define DT_LEARNED_ACCEPT_FIT 1e-4

typedef struct dt_hooi_args {
    int32_t ndim;
    int32_t dtype;              /* element type of A: DT_F32, DT_BF16 or DT_F64 */
    const void *A;              /* device */
    const float *A_f32;         /* optional FP32 copy (mixed precision polish / residual) */
    size_t shape[DT_MAX_NDIM];
    size_t ranks[DT_MAX_NDIM];
    int32_t max_iter;           /* sweeps (hooi_als.tpp:91) */
    int32_t inner_iter;         /* ALS_ inner iterations (hooi_als.tpp:9, default 5) */
    int32_t init;               /* DT_INIT_* */
    int32_t schedule;           /* DT_SCHED_* */
    double tol;                 /* stop early when |res_k - res_{k-1}| < tol; 0 = never */
    const double *U_init[DT_MAX_NDIM]; /* device FP64 I_n x R_n (GIVEN / QR) */
    uint64_t seed;              /* DT_INIT_RANDOM and the learned-init sketch */
    int32_t flags;              /* DT_FLAG_* */
    const void *model;          /* dt_mlp handle for DT_INIT_LEARNED */
} dt_hooi_args;

typedef struct dt_hooi_result {
    double *U[DT_MAX_NDIM];     /* device outputs FP64 I_n x R_n (NULL: not copied out) */
    double *G;                  /* device output core FP64, prod R_n (NULL: not copied) */
    double *residual_per_sweep; /* host array of max_iter entries (may be NULL) */
    int32_t sweeps;             /* sweeps run */
    int32_t init_used;          /* DT_INIT_* that produced the start factors */
    double A_norm;              /* ||A||_F */
    double G_norm;              /* ||G||_F after the last sweep */
    double residual;            /* sqrt(1 - ||G||^2/||A||^2) after the last sweep */
    double residual_direct;     /* ||A - G x_1 U_1 ... ||/||A|| (FP64), if requested */
    double init_ms;             /* device time of the start-factor computation */
    double sweeps_ms;           /* device time of all sweeps */
    double learned_first_res;   /* learned init: residual after its first sweep */
    double hosvd_first_res;     /* learned init: residual of the HOSVD comparison sweep */
} dt_hooi_result;

/* Algorithm::Tucker::HOOI_ALS (include/algorithm.hpp:12-14,
 * template/algorithm/tucker/hooi_als.tpp:64-123) on one GPU: create, max_iter
 * sweeps (early stop on tol), copy out, destroy.
 *
 * Device memory beside A: four intermediates of max_n (prod I / I_n) R_n
 * elements, the core, O(I_max R_max) factor buffers, and the mode Gram
 * S = Y_(n) Y_(n)ᵀ in FP64: 8 I_max^2 bytes (8 GiB at I_max = 32768; no
 * longer mode than about 190000 fits 288 GB).  A mode too long for S fails
 * in create with DT_ENOMEM naming S.  The reference's ALS_ never forms S
 * (W = ttt_except(Y, Y x_n Pᵀ), hooi_als.tpp:26-27); here every inner step
 * W = S P then costs an I x I x R product instead of two passes over Y. */
int dt_hooi_dense(dt_ctx *ctx, const dt_hooi_args *args, dt_hooi_result *res);

/* The same loop one sweep at a time (the reference's `for iter < max_iter`
 * body, hooi_als.tpp:91-115), for callers that time or steer sweeps.
 * dt_hooi_create computes ||A|| and the start factors (hooi_als.tpp:76-87);
 * dt_hooi_sweep runs one sweep (on the FP32 copy when fp32 != 0 and A is BF16)
 * and returns sqrt(1 - ||G||^2/||A||^2); dt_hooi_get copies U and G out (when
 * no sweep ran it forms G from the start factors). */
typedef struct dt_hooi dt_hooi;
int dt_hooi_create(dt_ctx *ctx, const dt_hooi_args *args, dt_hooi **out);
int dt_hooi_sweep(dt_hooi *h, int fp32, double *residual);
/* Residual of the current factors on A_f32 (or A): the core is formed again
 * from that tensor, in FP32 with the factor-copy correction, or (fp64 != 0)
 * with FP64 accumulation.  One pass over the tensor; the factors do not
 * change.  Lets a BF16 run report its fit on the FP32 copy without an FP32
 * sweep.  No reference counterpart (hooi_als.tpp:112-114 prints the residual
 * of the sweep's own core). */
int dt_hooi_fit(dt_hooi *h, int fp64, double *residual);
int dt_hooi_get(dt_hooi *h, dt_hooi_result *res);
int dt_hooi_destroy(dt_hooi *h);

#ifdef __cplusplus
}
#endif

#endif /* DIANA_HIP_H */
