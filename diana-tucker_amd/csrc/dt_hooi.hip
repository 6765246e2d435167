// Dense HOOI-ALS driver on one GPU.
//
// Reference: Algorithm::Tucker::HOOI_ALS (template/algorithm/tucker/hooi_als.tpp:64-123)
// and its per-mode update ALS_ (hooi_als.tpp:7-34).  Per sweep and mode n the
// reference forms Y = A x_{i != n} U_i^T by a chain of Function::ttm calls
// (hooi_als.tpp:94-101) and runs `inner_iter` (default 5) steps of
//     P = L (LᵀL)^{-1}  (P = L on the first step),  W = Y_(n) Y_(n)^T P,
//     G_R = Pᵀ W,  L = W G_R^{-1}
// where W is formed as ttt_except(Y, Y x_n Pᵀ, n) (hooi_als.tpp:21-30), then
// returns Q of QR(L) (hooi_als.tpp:32).  Here S = Y_(n) Y_(n)^T is formed once
// per update by the FP32 MFMA Gram kernel (this is the reference's own
// ALS_YYt_ formulation, hooi_als.tpp:36-62) and the inner steps run in FP64.
// The core is G = A x_1 U_1ᵀ ... x_N U_Nᵀ and the printed residual is
// sqrt(1 - ||G||^2 / ||A||^2) (hooi_als.tpp:105-114).
//
// Schedules: DT_SCHED_REFERENCE repeats the reference's chains (N+1 passes over
// A per sweep).  DT_SCHED_TREE splits the modes into halves L = {0..h-1},
// R = {h..N-1}: T_L = A x_{i in R} U_iᵀ serves every n in L, T_R = A x_{i in L}
// U_iᵀ (with the freshly updated U_L) serves every n in R, and the core is the
// last Y contracted on its own mode -- 2 passes over A, the same exact-arithmetic
// result because mode products along different modes commute.  Tree chains
// contract the highest mode first, so every step either has F = prod of
// uncontracted leading sizes (ttm_mk) or F = 1 (ttm_m0).
#include "dt_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

struct dt_hooi {
    dt_ctx *ctx = nullptr;
    dt_hooi_args args{};
    int N = 0;
    int64_t I[DT_MAX_NDIM] = {}, R[DT_MAX_NDIM] = {};
    int64_t total = 0;
    double normA2 = 0.0;       // ||A||^2 of A_f32 when given, else of A
    double normA2_sweep = 0.0; // ||A||^2 of the tensor the sweeps read (A itself)
    // factors and their kernel views
    double *U64[DT_MAX_NDIM] = {};
    float *U32[DT_MAX_NDIM] = {};
    float *M32[DT_MAX_NDIM] = {};
    uint16_t *Ub16[DT_MAX_NDIM] = {};
    double *M64[DT_MAX_NDIM] = {};
    // intermediates
    void *buf[4] = {};
    size_t buf_elems = 0;
    void *Gbuf = nullptr;  // core (work type)
    double *S = nullptr;   // In x In
    double *L = nullptr, *P = nullptr, *W = nullptr;
    double *Gm = nullptr, *Ginv = nullptr, *GR = nullptr, *GRinv = nullptr;
    double *work = nullptr;     // small FP64 scratch (R x R systems, QR passes)
    double *G64 = nullptr, *T64a = nullptr, *T64b = nullptr;  // core in FP64 + scratch
    double *Cfix[DT_MAX_NDIM] = {};  // (X_nᵀ X_n)^{-1} of the factor actually used
    double *Xfix = nullptr;      // FP64 copy of one kernel factor
    double *dscal = nullptr;     // device scalars: [0] = ||Ahat||^2
    int *dflag = nullptr;        // device breakdown flags
    void *fbuf[2] = {};          // FP64 full-pass buffers (DT_FLAG_FP64_FIT)
    int bf16_core_mode = -1;     // mode whose core-chain product used the bf16 factor
    bool args_reference_als = false;  // run ALS_ literally (inverse-normalised steps)
    double *alsws = nullptr;     // workspace of dt_als_update
    std::vector<void *> owned;
    int sweeps = 0;
    double last_res = -1.0;
    double G_norm = 0.0;
    double init_ms = 0.0, sweeps_ms = 0.0;
    int init_used = DT_INIT_GIVEN;
    int work_dt = DT_F32;  // element type of intermediates (F32, or F64 for F64 input)
};

static int hooi_alloc(dt_hooi *h, void **p, size_t bytes) {
    dt_ctx *ctx = h->ctx;
    if (bytes == 0) bytes = 256;
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return dt_fail(ctx, DT_ENOMEM, "hooi: hipMalloc(%zu) failed", bytes);
    }
    h->owned.push_back(*p);
    return DT_OK;
}

static int update_views(dt_hooi *h, int n) {
    dt_ctx *ctx = h->ctx;
    const bool f64 = h->args.dtype == DT_F64;
    return dt_f64_to_views(ctx, h->U64[n], (size_t)h->I[n], (size_t)h->R[n], f64 ? nullptr : h->U32[n],
                           f64 ? nullptr : h->M32[n], h->args.dtype == DT_BF16 ? h->Ub16[n] : nullptr,
                           f64 ? h->M64[n] : nullptr);
}

// dst = src x_n U_nᵀ, src shape sh (in/out: sh[n] becomes R_n).
static int contract(dt_hooi *h, const void *src, int src_dt, int64_t *sh, int n, void *dst) {
    dt_ctx *ctx = h->ctx;
    int64_t F = 1, B = 1;
    for (int i = 0; i < n; i++) F *= sh[i];
    for (int i = n + 1; i < h->N; i++) B *= sh[i];
    dt_factor_views fv{h->U32[n], h->M32[n], h->Ub16[n], h->M64[n]};
    const bool full = src == h->args.A || src == (const void *)h->args.A_f32;
    DT_TRY(dt_ttm_views(ctx, src, src_dt, F, sh[n], B, fv, h->R[n], dst, h->work_dt, full));
    sh[n] = h->R[n];
    return DT_OK;
}

// Contract src over `modes` in order using scratch buffers s0/s1 (never src).
static int chain(dt_hooi *h, const void *src, int src_dt, int64_t *sh, const std::vector<int> &modes,
                 void *s0, void *s1, const void **out, int *out_dt) {
    const void *cur = src;
    int cur_dt = src_dt;
    int which = 0;
    for (int m : modes) {
        void *dst = which == 0 ? s0 : s1;
        DT_TRY(contract(h, cur, cur_dt, sh, m, dst));
        cur = dst;
        cur_dt = h->work_dt;
        which ^= 1;
    }
    *out = cur;
    *out_dt = cur_dt;
    return DT_OK;
}

// ALS_ on Y (shape sh, mode n): U64[n] <- Q of QR(L after inner_iter steps).
// Small problems (I R^2 <= 2^20) run the R x R steps fused in one workgroup;
// larger ones use multi-workgroup GEMMs for the I x R products.  No step
// synchronises with the host; breakdowns set bits in h->dflag.
static int als_update(dt_hooi *h, const void *Y, int ydt, const int64_t *sh, int n) {
    dt_ctx *ctx = h->ctx;
    int64_t F = 1, B = 1;
    for (int i = 0; i < n; i++) F *= sh[i];
    for (int i = n + 1; i < h->N; i++) B *= sh[i];
    const size_t In = (size_t)sh[n], Rn = (size_t)h->R[n];
    // the fused R x R kernels hold their matrices in LDS (R <= 64)
    const bool small = (double)In * Rn * Rn <= (double)(1 << 20) && Rn <= 64;
    DT_TRY(dt_gram_internal(ctx, Y, ydt, F, (int64_t)In, B, h->S));
    if (!h->args_reference_als) {
        // span-equivalent normalised iteration, row blocks over many workgroups;
        // writes the FP64 factor and its kernel views
        const bool f64 = h->args.dtype == DT_F64;
        DT_TRY(dt_als_update(ctx, h->S, h->U64[n], (int)In, (int)Rn, h->args.inner_iter, h->alsws,
                             h->dflag, f64 ? nullptr : h->U32[n], f64 ? nullptr : h->M32[n],
                             h->args.dtype == DT_BF16 ? h->Ub16[n] : nullptr));
        if (f64) DT_TRY(update_views(h, n));
        return DT_OK;
    }
    (void)small;
    DT_TRY(dt_copy_f64(ctx, h->L, h->U64[n], In * Rn));
    for (int it = 0; it < h->args.inner_iter; it++) {
        const double *Pp = h->L;
        if (it > 0) {
            if (small) {
                DT_TRY(dt_als_prep(ctx, h->L, (int)In, (int)Rn, h->P, h->work, h->dflag));
            } else {
                DT_TRY(dt_gemm_f64_internal(ctx, 1, 0, Rn, Rn, In, 1.0, h->L, In, h->L, In, 0.0, h->Gm, Rn));
                DT_TRY(dt_inverse_flag(ctx, h->Gm, (int)Rn, h->Ginv, h->work, h->dflag, 1));
                DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, In, Rn, Rn, 1.0, h->L, In, h->Ginv, Rn, 0.0, h->P, In));
            }
            Pp = h->P;
        }
        DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, In, Rn, In, 1.0, h->S, In, Pp, In, 0.0, h->W, In));
        if (small) {
            DT_TRY(dt_als_finish(ctx, Pp, h->W, (int)In, (int)Rn, h->L, h->work, h->dflag));
        } else {
            DT_TRY(dt_gemm_f64_internal(ctx, 1, 0, Rn, Rn, In, 1.0, Pp, In, h->W, In, 0.0, h->GR, Rn));
            DT_TRY(dt_inverse_flag(ctx, h->GR, (int)Rn, h->GRinv, h->work, h->dflag, 2));
            DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, In, Rn, Rn, 1.0, h->W, In, h->GRinv, Rn, 0.0, h->L, In));
        }
    }
    DT_TRY(dt_copy_f64(ctx, h->U64[n], h->L, In * Rn));
    DT_TRY(dt_orth_dev(ctx, h->U64[n], In, Rn, h->work, h->dflag));
    return update_views(h, n);
}

// ||Ahat||^2 for the factors the kernels actually used.  With FP32 (or bf16)
// factor copies X_n, G = A x_n X_nᵀ and the projection of A onto
// span(X_1 (x) ... (x) X_N) has squared norm <G, G x_1 C_1 ... x_N C_N> with
// C_n = (X_nᵀ X_n)^{-1}; this removes the O(1e-8) bias that the rounding of
// the factors to FP32 puts into 1 - ||G||^2/||A||^2.  With DT_FLAG_FP64_FIT the
// core is instead re-formed from A with FP64 accumulation and FP64 factors.
static int ahat_sq_dev(dt_hooi *h, const void *A, int dt) {
    dt_ctx *ctx = h->ctx;
    const int N = h->N;
    int64_t csz = 1;
    for (int n = 0; n < N; n++) csz *= h->R[n];
    if (h->args.flags & DT_FLAG_FP64_FIT) {
        // FP64 chain from A: highest mode first (first step FP64-accumulating)
        int64_t sh[DT_MAX_NDIM];
        for (int i = 0; i < N; i++) sh[i] = h->I[i];
        const void *cur = A;
        int cdt = dt;
        int which = 0;
        for (int m = N - 1; m >= 0; m--) {
            int64_t F = 1, B = 1;
            for (int i = 0; i < m; i++) F *= sh[i];
            for (int i = m + 1; i < N; i++) B *= sh[i];
            double *dst = m == 0 ? h->G64 : (double *)h->fbuf[which];
            DT_TRY(dt_f64_to_views(ctx, h->U64[m], (size_t)h->I[m], (size_t)h->R[m], nullptr, nullptr,
                                   nullptr, h->Xfix));  // Xfix = M64 (R x I)
            DT_TRY(dt_ttm_f64acc(ctx, cur, cdt, F, sh[m], B, h->Xfix, h->R[m], dst));
            sh[m] = h->R[m];
            cur = dst;
            cdt = DT_F64;
            which ^= 1;
        }
        return dt_dot_f64_dev(ctx, h->G64, h->G64, (size_t)csz, h->dscal);
    }
    DT_TRY(dt_convert(ctx, h->G64, DT_F64, h->Gbuf, h->work_dt, (size_t)csz));
    if (h->work_dt == DT_F64) return dt_dot_f64_dev(ctx, h->G64, h->G64, (size_t)csz, h->dscal);
    // T = G x_n C_n for all n (generic FP64 mode products on the small core);
    // the small C_n come from one launch, straight from the FP32 / bf16 copies
    int64_t sh[DT_MAX_NDIM];
    for (int i = 0; i < N; i++) sh[i] = h->R[i];
    const double *cur = h->G64;
    auto small = [&](int n) { return h->R[n] <= 64; };  // the batched LDS kernel's orders
    dt_gram_inverse_batch gb{};
    for (int n = 0; n < N; n++) {
        if (!small(n)) continue;
        const bool b16 = n == h->bf16_core_mode && h->Ub16[n];
        gb.X[gb.count] = b16 ? (const void *)h->Ub16[n] : (const void *)h->U32[n];
        gb.dt[gb.count] = b16 ? DT_BF16 : DT_F32;
        gb.m[gb.count] = (int)h->I[n];
        gb.n[gb.count] = (int)h->R[n];
        gb.C[gb.count] = h->Cfix[n];
        gb.count++;
    }
    DT_TRY(dt_gram_inverse_batch_dev(ctx, gb, h->dflag));
    for (int n = 0; n < N; n++) {
        const size_t In = (size_t)h->I[n], Rn = (size_t)h->R[n];
        if (!small(n)) {
            if (n == h->bf16_core_mode && h->Ub16[n])
                DT_TRY(dt_convert(ctx, h->Xfix, DT_F64, h->Ub16[n], DT_BF16, In * Rn));
            else
                DT_TRY(dt_convert(ctx, h->Xfix, DT_F64, h->U32[n], DT_F32, In * Rn));
            DT_TRY(dt_gemm_f64_internal(ctx, 1, 0, Rn, Rn, In, 1.0, h->Xfix, In, h->Xfix, In, 0.0, h->Gm, Rn));
            DT_TRY(dt_inverse_flag(ctx, h->Gm, (int)Rn, h->Cfix[n], h->work, h->dflag, 8));
        }
        int64_t F = 1, B = 1;
        for (int i = 0; i < n; i++) F *= sh[i];
        for (int i = n + 1; i < N; i++) B *= sh[i];
        double *dst = (n % 2 == 0) ? h->T64a : h->T64b;
        dt_factor_views fv{nullptr, nullptr, nullptr, h->Cfix[n]};
        DT_TRY(dt_ttm_views(ctx, cur, DT_F64, F, (int64_t)Rn, B, fv, (int64_t)Rn, dst, DT_F64));
        cur = dst;
    }
    return dt_dot_f64_dev(ctx, cur, h->G64, (size_t)csz, h->dscal);
}

// sqrt(1 - ||G||^2 / normA2): normA2 must be the squared norm of the tensor the
// core was formed from
static int core_residual(dt_hooi *h, const void *A, int dt, double normA2, double *res) {
    dt_ctx *ctx = h->ctx;
    DT_TRY(ahat_sq_dev(h, A, dt));
    // one read-back per sweep: ||Ahat||^2 and the breakdown flags
    DT_HIP(hipMemcpyAsync(ctx->pinned, h->dscal, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    DT_HIP(hipMemcpyAsync((int *)(ctx->pinned + 1), h->dflag, sizeof(int), hipMemcpyDeviceToHost,
                          ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    const double g2 = ctx->pinned[0];
    const int flag = ((int *)(ctx->pinned + 1))[0];
    if (flag) {
        return dt_fail(ctx, DT_ENUMERIC, "hooi: numerical breakdown (flags 0x%x: %s%s%s%s)", flag,
                       (flag & 1) ? "singular L^T L; " : "", (flag & 2) ? "singular P^T S P; " : "",
                       (flag & 4) ? "rank-deficient factor in QR; " : "",
                       (flag & 8) ? "singular factor Gram; " : "");
    }
    h->G_norm = std::sqrt(g2);
    const double ratio = normA2 > 0.0 ? g2 / normA2 : 1.0;
    *res = std::sqrt(std::max(0.0, 1.0 - ratio));
    return DT_OK;
}

// One sweep over all modes with A (type dt) and the core at the end.
static int sweep_once(dt_hooi *h, const void *A, int dt) {
    const int N = h->N;
    int64_t sh[DT_MAX_NDIM];
    const void *Y;
    int ydt;
    h->bf16_core_mode = -1;
    if (h->args.schedule == DT_SCHED_TREE && N >= 2) {
        const int hm = N / 2;
        if (dt == DT_BF16) h->bf16_core_mode = hm - 1;  // first product of the T_R chain
        std::vector<int> left, right;
        for (int i = 0; i < hm; i++) left.push_back(i);
        for (int i = hm; i < N; i++) right.push_back(i);
        // Phase 1: T_L = A x_{R} (last mode first), Y_n for n in L
        for (int i = 0; i < N; i++) sh[i] = h->I[i];
        std::vector<int> rdesc(right.rbegin(), right.rend());
        const void *TL;
        int tldt;
        int tfused = DT_EUNSUPPORTED;
        if (rdesc.size() >= 2 && h->work_dt == DT_F32 && dt == DT_F32) {
            // modes N-1 and N-2 in one pass: no intermediate of size |A| R_{N-1} / I_{N-1}
            const int m2 = N - 2, m3 = N - 1;
            int64_t F = 1;
            for (int i = 0; i < m2; i++) F *= sh[i];
            tfused = dt_ttm2_trail(h->ctx, (const float *)A, dt, F, sh[m2], sh[m3], h->M32[m2], h->M32[m3],
                                   h->R[m2], h->R[m3], (float *)h->buf[0]);
            if (tfused != DT_OK && tfused != DT_EUNSUPPORTED) return tfused;
            if (tfused == DT_OK) {
                sh[m2] = h->R[m2];
                sh[m3] = h->R[m3];
                std::vector<int> rest(rdesc.begin() + 2, rdesc.end());
                DT_TRY(chain(h, h->buf[0], DT_F32, sh, rest, h->buf[1], h->buf[0], &TL, &tldt));
            }
        }
        if (tfused != DT_OK) DT_TRY(chain(h, A, dt, sh, rdesc, h->buf[0], h->buf[1], &TL, &tldt));
        int64_t shL[DT_MAX_NDIM];
        for (int n : left) {
            std::copy(sh, sh + N, shL);
            std::vector<int> rest;
            for (int i = hm - 1; i >= 0; i--)
                if (i != n) rest.push_back(i);
            DT_TRY(chain(h, TL, tldt, shL, rest, h->buf[2], h->buf[3], &Y, &ydt));
            DT_TRY(als_update(h, Y, ydt, shL, n));
        }
        // Phase 2: T_R = A x_{L} (ascending), Y_n for n in R
        for (int i = 0; i < N; i++) sh[i] = h->I[i];
        const void *TR;
        int trdt;
        std::vector<int> ldesc(left.rbegin(), left.rend());
        int fused = DT_EUNSUPPORTED;
        if (hm >= 2 && h->work_dt == DT_F32 && (dt == DT_F32 || dt == DT_BF16)) {
            // modes 0 and 1 in one pass: no intermediate of size |A| R_1 / I_1
            int64_t B = 1;
            for (int i = 2; i < N; i++) B *= sh[i];
            const void *U0 = dt == DT_BF16 ? (const void *)h->Ub16[0] : (const void *)h->U32[0];
            fused = dt_ttm2_lead(h->ctx, A, dt, sh[0], sh[1], B, U0, h->M32[1], h->R[0], h->R[1],
                                 (float *)h->buf[0]);
            if (fused != DT_OK && fused != DT_EUNSUPPORTED) return fused;
        }
        if (fused == DT_OK) {
            if (dt == DT_BF16) h->bf16_core_mode = 0;
            sh[0] = h->R[0];
            sh[1] = h->R[1];
            std::vector<int> rest(ldesc.begin(), ldesc.end() - 2);  // modes hm-1 .. 2
            DT_TRY(chain(h, h->buf[0], DT_F32, sh, rest, h->buf[1], h->buf[0], &TR, &trdt));
        } else {
            DT_TRY(chain(h, A, dt, sh, ldesc, h->buf[0], h->buf[1], &TR, &trdt));
        }
        int64_t shR[DT_MAX_NDIM];
        const void *Ylast = nullptr;
        int ylastdt = 0;
        for (int n : right) {
            std::copy(sh, sh + N, shR);
            std::vector<int> rest;
            for (int i = N - 1; i >= hm; i--)
                if (i != n) rest.push_back(i);
            DT_TRY(chain(h, TR, trdt, shR, rest, h->buf[2], h->buf[3], &Y, &ydt));
            DT_TRY(als_update(h, Y, ydt, shR, n));
            Ylast = Y;
            ylastdt = ydt;
        }
        // core from the last Y (mode N-1 still uncontracted)
        DT_TRY(contract(h, Ylast, ylastdt, shR, N - 1, h->Gbuf));
        return DT_OK;
    }
    // reference schedule
    if (dt == DT_BF16) h->bf16_core_mode = 0;
    for (int n = 0; n < N; n++) {
        for (int i = 0; i < N; i++) sh[i] = h->I[i];
        std::vector<int> modes;
        for (int i = 0; i < N; i++)
            if (i != n) modes.push_back(i);
        DT_TRY(chain(h, A, dt, sh, modes, h->buf[0], h->buf[1], &Y, &ydt));
        DT_TRY(als_update(h, Y, ydt, sh, n));
    }
    for (int i = 0; i < N; i++) sh[i] = h->I[i];
    std::vector<int> all;
    for (int i = 0; i < N - 1; i++) all.push_back(i);
    const void *T;
    int tdt;
    DT_TRY(chain(h, A, dt, sh, all, h->buf[0], h->buf[1], &T, &tdt));
    return contract(h, T, tdt, sh, N - 1, h->Gbuf);
}

// Core from the current factors (reference chain) of tensor src (type sdt).
static int core_from(dt_hooi *h, const void *src, int sdt) {
    h->bf16_core_mode = sdt == DT_BF16 ? 0 : -1;
    int64_t sh[DT_MAX_NDIM];
    for (int i = 0; i < h->N; i++) sh[i] = h->I[i];
    std::vector<int> all;
    for (int i = 0; i < h->N - 1; i++) all.push_back(i);
    const void *T;
    int tdt;
    DT_TRY(chain(h, src, sdt, sh, all, h->buf[0], h->buf[1], &T, &tdt));
    return contract(h, T, tdt, sh, h->N - 1, h->Gbuf);
}

// Core from the current factors, for max_iter == 0.
static int core_only(dt_hooi *h) { return core_from(h, h->args.A, h->args.dtype); }

int dt_learned_init(dt_ctx *ctx, const void *model, const void *A, int dtA, int N, const int64_t *I,
                    const int64_t *R, int sketch, uint64_t seed, double *const *U64);

static int init_factors(dt_hooi *h) {
    dt_ctx *ctx = h->ctx;
    const dt_hooi_args &a = h->args;
    const int N = h->N;
    h->init_used = a.init;
    if (a.init == DT_INIT_GIVEN || a.init == DT_INIT_QR) {
        for (int n = 0; n < N; n++) {
            DT_CHECK_ARG(a.U_init[n], "hooi: U_init[%d] missing", n);
            if (a.init == DT_INIT_GIVEN)
                DT_TRY(dt_copy_f64(ctx, h->U64[n], a.U_init[n], (size_t)(h->I[n] * h->R[n])));
            else
                DT_TRY(dt_qr_f64(ctx, h->U64[n], nullptr, a.U_init[n], (size_t)h->I[n], (size_t)h->R[n]));
        }
    } else if (a.init == DT_INIT_RANDOM) {
        for (int n = 0; n < N; n++) {
            DT_TRY(dt_fill_normal_f64(ctx, h->U64[n], (size_t)(h->I[n] * h->R[n]), a.seed, 100 + (uint64_t)n));
            DT_TRY(dt_qr_f64(ctx, h->U64[n], nullptr, h->U64[n], (size_t)h->I[n], (size_t)h->R[n]));
        }
    } else if (a.init == DT_INIT_HOSVD) {
        const void *src = a.A_f32 ? (const void *)a.A_f32 : a.A;
        const int sdt = a.A_f32 ? DT_F32 : a.dtype;
        for (int n = 0; n < N; n++) {
            int64_t F = 1, B = 1;
            for (int i = 0; i < n; i++) F *= h->I[i];
            for (int i = n + 1; i < N; i++) B *= h->I[i];
            DT_TRY(dt_gram_internal(ctx, src, sdt, F, h->I[n], B, h->S));
            DT_TRY(dt_side_kick(ctx));  // the eigen step is latency-bound: the norms overlap it
            DT_TRY(dt_syev_top_f64(ctx, h->S, (size_t)h->I[n], (size_t)h->R[n], h->U64[n], nullptr));
        }
    } else if (a.init == DT_INIT_LEARNED) {
        DT_TRY(dt_learned_init(ctx, a.model, a.A, a.dtype, N, h->I, h->R,
                               (a.flags & DT_FLAG_SKETCH_KRON) ? DT_SKETCH_KRONECKER : DT_SKETCH_GAUSSIAN, a.seed,
                               h->U64));
    } else {
        return dt_fail(ctx, DT_EINVAL, "hooi: unknown init %d", a.init);
    }
    for (int n = 0; n < N; n++) DT_TRY(update_views(h, n));
    return DT_OK;
}

extern "C" {

int dt_hooi_destroy(dt_hooi *h) {
    if (!h) return DT_OK;
    hipStreamSynchronize(h->ctx->stream);
    for (void *p : h->owned) hipFree(p);
    delete h;
    return DT_OK;
}

int dt_hooi_create(dt_ctx *ctx, const dt_hooi_args *args, dt_hooi **out) {
    DT_CHECK_ARG(ctx && args && out && args->A, "dt_hooi_create: null argument");
    *out = nullptr;
    const int N = args->ndim;
    DT_CHECK_ARG(N >= 2 && N <= DT_MAX_NDIM, "hooi: ndim %d not in [2, %d]", N, DT_MAX_NDIM);
    DT_CHECK_ARG(args->dtype == DT_F32 || args->dtype == DT_BF16 || args->dtype == DT_F64,
                 "hooi: dtype %d", args->dtype);
    DT_CHECK_ARG(args->inner_iter >= 0 && args->max_iter >= 0, "hooi: negative iteration count");
    dt_hooi *h = new dt_hooi();
    h->ctx = ctx;
    h->args = *args;
    h->N = N;
    h->work_dt = args->dtype == DT_F64 ? DT_F64 : DT_F32;
    h->total = 1;
    int64_t maxI = 0, maxR = 0;
    for (int n = 0; n < N; n++) {
        h->I[n] = (int64_t)args->shape[n];
        h->R[n] = (int64_t)args->ranks[n];
        if (h->I[n] < 1 || h->R[n] < 1 || h->R[n] > h->I[n]) {
            delete h;
            return dt_fail(ctx, DT_EINVAL, "hooi: mode %d has size %zu and rank %zu (need 1 <= R <= I)",
                           n, args->shape[n], args->ranks[n]);
        }
        h->total *= h->I[n];
        maxI = std::max(maxI, h->I[n]);
        maxR = std::max(maxR, h->R[n]);
    }
    for (int n = 0; n < N; n++) {
        // Y_(n) has prod_{m != n} R_m columns: a larger R_n leaves its Gram singular
        int64_t k = 1;
        for (int m = 0; m < N; m++)
            if (m != n) k *= h->R[m];
        if (h->R[n] > k) {
            delete h;
            return dt_fail(ctx, DT_EINVAL, "hooi: rank %zu of mode %d above the product %lld of the other ranks",
                           args->ranks[n], n, (long long)k);
        }
    }
    const size_t wsz = h->work_dt == DT_F64 ? 8 : 4;
    int st = DT_OK;
    // largest intermediate: A with one mode contracted
    size_t maxe = 0;
    for (int n = 0; n < N; n++) maxe = std::max(maxe, (size_t)(h->total / h->I[n] * h->R[n]));
    h->buf_elems = maxe;
    for (int b = 0; b < 4 && st == DT_OK; b++) st = hooi_alloc(h, &h->buf[b], maxe * wsz);
    int64_t csz = 1;
    for (int n = 0; n < N; n++) csz *= h->R[n];
    csz = std::max<int64_t>(csz, 64);
    if (st == DT_OK) st = hooi_alloc(h, &h->Gbuf, (size_t)csz * wsz);
    if (st == DT_OK && hooi_alloc(h, (void **)&h->S, (size_t)(maxI * maxI) * 8) != DT_OK)
        st = dt_fail(ctx, DT_ENOMEM, "hooi: the FP64 mode Gram S of the longest mode (%lld) needs %.3g GB",
                     (long long)maxI, (double)maxI * (double)maxI * 8e-9);
    const size_t ir = (size_t)(maxI * maxR) * 8, rr = (size_t)(maxR * maxR) * 8;
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->L, ir);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->P, ir);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->W, ir);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->Gm, rr);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->Ginv, rr);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->GR, rr);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->GRinv, rr);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->work, std::max(4 * rr, 2 * rr + ir) + 1024);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->G64, (size_t)csz * 8);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->T64a, (size_t)csz * 8);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->T64b, (size_t)csz * 8);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->Xfix, ir);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->dscal, 256);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->dflag, 256);
    if (st == DT_OK) st = hooi_alloc(h, (void **)&h->alsws, dt_als_ws_doubles(maxI, maxR) * 8);
    h->args_reference_als = (args->flags & DT_FLAG_REFERENCE_ALS) != 0;
    for (int n = 0; n < N && st == DT_OK; n++) st = hooi_alloc(h, (void **)&h->Cfix[n], rr);
    if (st == DT_OK && (args->flags & DT_FLAG_FP64_FIT)) {
        st = hooi_alloc(h, &h->fbuf[0], maxe * 8);
        if (st == DT_OK) st = hooi_alloc(h, &h->fbuf[1], maxe * 8);
    }
    for (int n = 0; n < N && st == DT_OK; n++) {
        const size_t e = (size_t)(h->I[n] * h->R[n]);
        st = hooi_alloc(h, (void **)&h->U64[n], e * 8);
        if (st == DT_OK && h->work_dt == DT_F32) st = hooi_alloc(h, (void **)&h->U32[n], e * 4);
        if (st == DT_OK && h->work_dt == DT_F32) st = hooi_alloc(h, (void **)&h->M32[n], e * 4);
        if (st == DT_OK && args->dtype == DT_BF16) st = hooi_alloc(h, (void **)&h->Ub16[n], e * 2);
        if (st == DT_OK && h->work_dt == DT_F64) st = hooi_alloc(h, (void **)&h->M64[n], e * 8);
    }
    if (st != DT_OK) {
        dt_hooi_destroy(h);
        return st;
    }
    // ||A|| of the FP32 copy when one is given (mixed precision), else of A,
    // and ||A||^2 of the BF16 tensor the sweeps read: side work that the start
    // launches where its eigen step leaves HBM idle (dt_side_kick), joined
    // before the start returns
    const void *nsrc[2] = {args->A_f32 ? (const void *)args->A_f32 : args->A, args->A};
    const int ndt[2] = {args->A_f32 ? DT_F32 : args->dtype, args->dtype};
    const size_t nlen[2] = {(size_t)h->total, (size_t)h->total};
    const int nsum = (args->A_f32 && args->dtype != DT_F32) ? 2 : 1;
    st = dt_side_sumsq_setup(ctx, nsum, nsrc, ndt, nlen);
    if (st == DT_OK) {
        hipEventRecord(ctx->t0, ctx->stream);
        st = init_factors(h);
        hipEventRecord(ctx->t1, ctx->stream);
        hipEventSynchronize(ctx->t1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, ctx->t0, ctx->t1);
        h->init_ms = ms;
    }
    double norms[2] = {0.0, 0.0};
    const int jst = dt_side_join(ctx, st == DT_OK ? norms : nullptr);
    if (st == DT_OK) st = jst;
    if (st == DT_OK) {
        h->normA2 = norms[0];
        h->normA2_sweep = nsum == 2 ? norms[1] : norms[0];
        if (!(h->normA2 > 0.0 && std::isfinite(h->normA2)))
            st = dt_fail(ctx, DT_ENUMERIC, "hooi: ||A|| is zero or not finite");
    }
    if (st != DT_OK) {
        dt_hooi_destroy(h);
        return st;
    }
    *out = h;
    return DT_OK;
}

// One sweep; residual sqrt(1 - ||G||^2/||A||^2) after it.  With fp32 != 0 and
// a BF16 tensor the sweep runs on the FP32 copy instead.
int dt_hooi_sweep(dt_hooi *h, int fp32, double *residual) {
    dt_ctx *ctx = h->ctx;
    const void *A = h->args.A;
    int dt = h->args.dtype;
    if (fp32 && dt == DT_BF16) {
        DT_CHECK_ARG(h->args.A_f32, "hooi: FP32 sweep needs A_f32");
        A = h->args.A_f32;
        dt = DT_F32;
    }
    hipEvent_t e0 = dt_get_event(ctx), e1 = dt_get_event(ctx);
    hipEventRecord(e0, ctx->stream);
    int st = dt_zero(ctx, h->dflag, sizeof(int));
    if (st == DT_OK) st = sweep_once(h, A, dt);
    double res = 0.0;
    const void *fitA = h->args.A_f32 ? (const void *)h->args.A_f32 : A;
    const int fitdt = h->args.A_f32 ? DT_F32 : dt;
    // with DT_FLAG_FP64_FIT the core is formed again from fitA; otherwise it is
    // the sweep's own, from the tensor the sweep read
    const bool refit = (h->args.flags & DT_FLAG_FP64_FIT) != 0;
    const double n2 = refit || A != h->args.A ? h->normA2 : h->normA2_sweep;
    if (st == DT_OK) st = core_residual(h, fitA, fitdt, n2, &res);
    hipEventRecord(e1, ctx->stream);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    ctx->event_pool.push_back(e0);
    ctx->event_pool.push_back(e1);
    if (st != DT_OK) return st;
    if (!std::isfinite(res)) return dt_fail(ctx, DT_ENUMERIC, "hooi: non-finite residual");
    h->sweeps_ms += ms;
    h->sweeps += 1;
    h->last_res = res;
    if (residual) *residual = res;
    return DT_OK;
}

int dt_hooi_fit(dt_hooi *h, int fp64, double *residual) {
    dt_ctx *ctx = h->ctx;
    DT_CHECK_ARG(residual, "dt_hooi_fit: null residual");
    const void *fitA = h->args.A_f32 ? (const void *)h->args.A_f32 : h->args.A;
    const int fitdt = h->args.A_f32 ? DT_F32 : h->args.dtype;
    DT_TRY(dt_zero(ctx, h->dflag, sizeof(int)));
    if (fp64 && !h->fbuf[0]) {
        // a session created without DT_FLAG_FP64_FIT has no FP64 chain buffers yet
        DT_TRY(hooi_alloc(h, &h->fbuf[0], h->buf_elems * 8));
        DT_TRY(hooi_alloc(h, &h->fbuf[1], h->buf_elems * 8));
    }
    const int saved = h->args.flags;
    if (fp64) {
        h->args.flags |= DT_FLAG_FP64_FIT;
    } else {
        h->args.flags &= ~DT_FLAG_FP64_FIT;
        DT_TRY(core_from(h, fitA, fitdt));
    }
    const int st = core_residual(h, fitA, fitdt, h->normA2, residual);
    h->args.flags = saved;
    if (st == DT_OK) h->last_res = *residual;
    return st;
}

int dt_hooi_get(dt_hooi *h, dt_hooi_result *res) {
    dt_ctx *ctx = h->ctx;
    if (h->sweeps == 0) {
        DT_TRY(dt_zero(ctx, h->dflag, sizeof(int)));
        DT_TRY(core_only(h));
        double r = 0.0;
        const void *fitA = h->args.A_f32 ? (const void *)h->args.A_f32 : h->args.A;
        const double n2 = (h->args.flags & DT_FLAG_FP64_FIT) ? h->normA2 : h->normA2_sweep;
        DT_TRY(core_residual(h, fitA, h->args.A_f32 ? DT_F32 : h->args.dtype, n2, &r));
        h->last_res = r;
    }
    int64_t csz = 1;
    for (int n = 0; n < h->N; n++) csz *= h->R[n];
    for (int n = 0; n < h->N; n++)
        if (res->U[n]) DT_TRY(dt_copy_f64(ctx, res->U[n], h->U64[n], (size_t)(h->I[n] * h->R[n])));
    if (res->G) DT_TRY(dt_convert(ctx, res->G, DT_F64, h->Gbuf, h->work_dt, (size_t)csz));
    res->sweeps = h->sweeps;
    res->init_used = h->init_used;
    res->A_norm = std::sqrt(h->normA2);
    res->G_norm = h->G_norm;
    res->residual = h->last_res;
    res->init_ms = h->init_ms;
    res->sweeps_ms = h->sweeps_ms;
    DT_HIP(hipStreamSynchronize(ctx->stream));
    return DT_OK;
}

int dt_hooi_dense(dt_ctx *ctx, const dt_hooi_args *args, dt_hooi_result *res) {
    DT_CHECK_ARG(ctx && args && res, "dt_hooi_dense: null argument");
    dt_hooi *h = nullptr;
    DT_TRY(dt_hooi_create(ctx, args, &h));
    int st = DT_OK;
    double prev = -1.0;
    const bool polish = (args->flags & DT_FLAG_FP32_POLISH) && args->dtype == DT_BF16 && args->A_f32;
    int first = 0;
    if (args->init == DT_INIT_LEARNED && args->max_iter > 0) {
        // Learned start, then its first sweep.  Fallback to the HOSVD start when
        // that sweep's fit is not a number in (0, 1], and, with
        // DT_FLAG_LEARNED_CHECK, when it is worse than the first sweep from the
        // HOSVD start.
        double rl = 0.0, rl2 = -1.0;
        st = dt_hooi_sweep(h, polish && args->max_iter == 1, &rl);
        // acceptance: one more sweep moves the fit by less than the fit
        // tolerance, i.e. HOOI from the learned start has converged to that
        // tolerance after its first sweep; otherwise the HOSVD start is taken.
        // With max_iter >= 2 that sweep is the run's own second sweep.
        if (st == DT_OK && (args->flags & DT_FLAG_LEARNED_ACCEPT)) st = dt_hooi_sweep(h, 0, &rl2);
        if (st != DT_OK) {
            dt_hooi_destroy(h);
            return st;
        }
        res->learned_first_res = rl;
        res->hosvd_first_res = -1.0;
        bool fallback = !(rl >= 0.0 && rl < 1.0);
        if ((args->flags & DT_FLAG_LEARNED_ACCEPT) && !(std::fabs(rl2 - rl) < DT_LEARNED_ACCEPT_FIT)) fallback = true;
        dt_hooi *hh = nullptr;
        double rh = 0.0;
        if (fallback || (args->flags & DT_FLAG_LEARNED_CHECK)) {
            dt_hooi_args a2 = *args;
            a2.init = DT_INIT_HOSVD;
            st = dt_hooi_create(ctx, &a2, &hh);
            if (st == DT_OK) st = dt_hooi_sweep(hh, polish && args->max_iter == 1, &rh);
            if (st != DT_OK) {
                dt_hooi_destroy(hh);
                dt_hooi_destroy(h);
                return st;
            }
            res->hosvd_first_res = rh;
            if (rh < rl) fallback = true;
        }
        if (fallback) {
            dt_hooi_destroy(h);
            h = hh;
            h->init_used = DT_INIT_HOSVD;
        } else if (hh) {
            dt_hooi_destroy(hh);
        }
        const double r = fallback ? rh : rl;
        if (res->residual_per_sweep) res->residual_per_sweep[0] = r;
        prev = r;
        first = 1;
        if (!fallback && rl2 >= 0.0 && args->max_iter >= 2) {
            // the acceptance sweep is the run's second sweep
            if (res->residual_per_sweep) res->residual_per_sweep[1] = rl2;
            prev = rl2;
            first = 2;
        }
    }
    for (int it = first; it < args->max_iter; it++) {
        const bool last = it == args->max_iter - 1;
        double r = 0.0;
        st = dt_hooi_sweep(h, polish && last, &r);
        if (st != DT_OK) break;
        if (res->residual_per_sweep) res->residual_per_sweep[it] = r;
        if (args->tol > 0.0 && prev >= 0.0 && std::fabs(r - prev) < args->tol) {
            if (polish && !last) {
                st = dt_hooi_sweep(h, 1, &r);
                if (st == DT_OK && res->residual_per_sweep && it + 1 < args->max_iter)
                    res->residual_per_sweep[it + 1] = r;
            }
            break;
        }
        prev = r;
    }
    if (st == DT_OK) st = dt_hooi_get(h, res);
    dt_hooi_destroy(h);
    return st;
}

}  // extern "C"
