// Per-mode factor update of dense HOOI, many workgroups per step.
//
// Reference ALS_ (template/algorithm/tucker/hooi_als.tpp:7-34), with
// S = Y_(n) Y_(n)^T, runs `inner` steps of
//     P = L (LᵀL)^{-1},  W = S P,  L = W (Pᵀ W)^{-1}
// and returns Q of QR(L).  Every step multiplies the current basis by S and
// re-normalises it by an R x R matrix, so span(L_k) = span(S^k L_0) and the
// returned factor is an orthonormal basis of span(S^inner U).  This file
// computes the same span with a normalisation that is cheaper and better
// conditioned on the GPU:
//     Q_0 = U,   Y_{k+1} = S Q_k = (S Y_k) R_k^{-1},   R_k = chol(Y_kᵀ Y_k)
// (Cholesky-QR of every iterate, never formed explicitly: the R_k^{-1} factor
// is applied to the rows of S Y_k).  A last Cholesky-QR pass on Q_inner makes
// the factor orthonormal to FP64 precision.  Equal spans in exact arithmetic;
// the parity tests compare fits and principal angles with the oracle, which
// runs the reference's own formulation.
//
//   als_rows_kernel  : RB rows per workgroup: T = S[rows,:] Y (S symmetric, so a
//                      row of S is a contiguous column), out = T Rinv, plus the
//                      workgroup's partial Gram outᵀ out.  Optionally writes the
//                      FP32 / bf16 / transposed kernel views of the result.
//   als_chol_kernel  : one workgroup: G = sum of the partials (fixed order),
//                      optional shift 11 (I R + R(R+1)) u tr(G) (Fukaya et al.
//                      2020), G = RᵀR, writes R^{-1}.
#include "dt_internal.hpp"

#include <cmath>

constexpr int ALS_RB = 16;      // rows per workgroup
constexpr int ALS_THREADS = 256;
constexpr int ALS_CH = 32;      // k-chunk staged in LDS
constexpr int ALS_MAXR = 64;

// mode 0: out = (S[rows,:] Y) Rinv      (S: I x I symmetric)
// mode 1: out = Y[rows,:] Rinv          (no S)
// Rinv == nullptr means identity.  part (R x R per workgroup) receives outᵀout
// of this workgroup's rows when non-null.
template <int RMAX>
__global__ __launch_bounds__(ALS_THREADS) void als_rows_kernel(
    int mode, const double *__restrict__ S, const double *__restrict__ Y, int I, int R,
    const double *__restrict__ Rinv, double *__restrict__ out, double *__restrict__ part,
    float *__restrict__ U32, float *__restrict__ M32, uint16_t *__restrict__ Ub16) {
    __shared__ double Ss[ALS_CH][ALS_RB + 1];     // S[t0+c, row0+r]
    __shared__ double Ys[ALS_CH][RMAX + 1];       // Y[t0+c, j]
    __shared__ double T[ALS_RB][RMAX + 1];        // rows x R
    __shared__ double Ri[RMAX][RMAX + 1];
    const int tid = threadIdx.x;
    const int row0 = blockIdx.x * ALS_RB;
    const int nrow = (I - row0) < ALS_RB ? (I - row0) : ALS_RB;

    for (int e = tid; e < R * R; e += ALS_THREADS) {
        const int i = e % R, j = e / R;
        Ri[i][j] = Rinv ? Rinv[e] : (i == j ? 1.0 : 0.0);
    }
    // outputs owned by this thread: (r, j) = (o % RB, o / RB), o = tid + q*256
    constexpr int NQ = (ALS_RB * RMAX + ALS_THREADS - 1) / ALS_THREADS;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) acc[q] = 0.0;

    if (mode == 0) {
        for (int t0 = 0; t0 < I; t0 += ALS_CH) {
            const int ch = (I - t0) < ALS_CH ? (I - t0) : ALS_CH;
            for (int e = tid; e < ALS_CH * ALS_RB; e += ALS_THREADS) {
                const int c = e % ALS_CH, r = e / ALS_CH;
                Ss[c][r] = (c < ch && r < nrow) ? S[(t0 + c) + (size_t)I * (row0 + r)] : 0.0;
            }
            for (int e = tid; e < ALS_CH * R; e += ALS_THREADS) {
                const int c = e % ALS_CH, j = e / ALS_CH;
                Ys[c][j] = c < ch ? Y[(t0 + c) + (size_t)I * j] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int o = tid + q * ALS_THREADS;
                const int r = o % ALS_RB, j = o / ALS_RB;
                if (j < R) {
                    double s = acc[q];
#pragma unroll 8
                    for (int c = 0; c < ALS_CH; c++) s = fma(Ss[c][r], Ys[c][j], s);
                    acc[q] = s;
                }
            }
            __syncthreads();
        }
    } else {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int o = tid + q * ALS_THREADS;
            const int r = o % ALS_RB, j = o / ALS_RB;
            if (j < R && r < nrow) acc[q] = Y[(row0 + r) + (size_t)I * j];
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int o = tid + q * ALS_THREADS;
        const int r = o % ALS_RB, j = o / ALS_RB;
        if (j < R) T[r][j] = acc[q];
    }
    __syncthreads();
    // out = T Rinv (Rinv upper triangular or identity; full product is cheap)
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int o = tid + q * ALS_THREADS;
        const int r = o % ALS_RB, j = o / ALS_RB;
        if (j < R) {
            double s = 0.0;
            for (int m = 0; m < R; m++) s = fma(T[r][m], Ri[m][j], s);
            acc[q] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int o = tid + q * ALS_THREADS;
        const int r = o % ALS_RB, j = o / ALS_RB;
        if (j < R) {
            T[r][j] = r < nrow ? acc[q] : 0.0;
            if (r < nrow) {
                const size_t gi = (size_t)(row0 + r) + (size_t)I * j;
                out[gi] = acc[q];
                if (U32) U32[gi] = (float)acc[q];
                if (Ub16) Ub16[gi] = f32_to_bf16((float)acc[q]);
                if (M32) M32[j + (size_t)R * (row0 + r)] = (float)acc[q];
            }
        }
    }
    __syncthreads();
    if (part) {
        double *p = part + (size_t)blockIdx.x * R * R;
        for (int e = tid; e < R * R; e += ALS_THREADS) {
            const int i = e % R, j = e / R;
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < ALS_RB; r++) s = fma(T[r][i], T[r][j], s);
            p[e] = s;
        }
    }
}

// One workgroup: G = sum_w part[w]; optional shift; Rinv = chol(G)^{-1}.
__global__ __launch_bounds__(ALS_THREADS) void als_chol_kernel(const double *__restrict__ part,
                                                               int nparts, int R, int I,
                                                               int shifted,
                                                               double *__restrict__ Rinv_out,
                                                               int *__restrict__ flag) {
    __shared__ double G[ALS_MAXR][ALS_MAXR + 1];
    __shared__ double shift_s, rjj_s;
    __shared__ int bad_s;
    const int tid = threadIdx.x;
    for (int e = tid; e < R * R; e += ALS_THREADS) {
        double s = 0.0;
        for (int w = 0; w < nparts; w++) s += part[(size_t)w * R * R + e];
        G[e % R][e / R] = s;
    }
    __syncthreads();
    if (tid == 0) {
        double shift = 0.0;
        if (shifted) {
            double tr = 0.0;
            for (int i = 0; i < R; i++) tr += G[i][i];
            shift = 11.0 * ((double)I * R + (double)R * (R + 1)) * 1.1102230246251565e-16 * tr;
        }
        shift_s = shift;
        bad_s = 0;
    }
    __syncthreads();
    if (tid < R) G[tid][tid] += shift_s;
    __syncthreads();
    for (int j = 0; j < R; j++) {
        if (tid == 0) {
            const double d = G[j][j];
            if (!(d > 0.0) || !isfinite(d)) {
                bad_s = j + 1;
                rjj_s = 1.0;
            } else {
                rjj_s = sqrt(d);
                G[j][j] = rjj_s;
            }
        }
        __syncthreads();
        if (bad_s) {
            if (tid == 0) atomicOr(flag, 4);
            return;
        }
        const double rjj = rjj_s;
        for (int k = j + 1 + tid; k < R; k += ALS_THREADS) G[j][k] /= rjj;
        __syncthreads();
        const int mm = R - j - 1;
        for (int idx = tid; idx < mm * mm; idx += ALS_THREADS) {
            const int i = j + 1 + idx % mm, k = j + 1 + idx / mm;
            if (i <= k) G[i][k] -= G[j][i] * G[j][k];
        }
        __syncthreads();
    }
    // R^{-1}: column j by thread j, back substitution straight into the
    // output (each thread only reads back its own column).
    for (int e = tid; e < R * R; e += ALS_THREADS)
        if (e % R > e / R) Rinv_out[e] = 0.0;
    if (tid < R) {
        const int j = tid;
        double *x = Rinv_out + (size_t)R * j;
        for (int i = j; i >= 0; i--) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int k = i + 1; k <= j; k++) s -= G[i][k] * x[k];
            x[i] = s / G[i][i];
        }
    }
}

template <int RMAX>
static int launch_rows(dt_ctx *ctx, int mode, const double *S, const double *Y, int I, int R,
                       const double *Rinv, double *out, double *part, float *U32, float *M32,
                       uint16_t *Ub16) {
    const int nb = (I + ALS_RB - 1) / ALS_RB;
    DT_LAUNCH("als_rows", als_rows_kernel<RMAX>, dim3(nb), dim3(ALS_THREADS), 0, mode, S, Y, I, R, Rinv,
              out, part, U32, M32, Ub16);
    return DT_OK;
}

static int rows(dt_ctx *ctx, int mode, const double *S, const double *Y, int I, int R,
                const double *Rinv, double *out, double *part, float *U32 = nullptr,
                float *M32 = nullptr, uint16_t *Ub16 = nullptr) {
    if (R <= 16) return launch_rows<16>(ctx, mode, S, Y, I, R, Rinv, out, part, U32, M32, Ub16);
    if (R <= 32) return launch_rows<32>(ctx, mode, S, Y, I, R, Rinv, out, part, U32, M32, Ub16);
    return launch_rows<64>(ctx, mode, S, Y, I, R, Rinv, out, part, U32, M32, Ub16);
}

// Workspace needs: part (ceil(I/16) R^2) + 2 Rinv (R^2) + 2 panels (I R) doubles.
size_t dt_als_ws_doubles(size_t I, size_t R) {
    return ((I + ALS_RB - 1) / ALS_RB) * R * R + 2 * R * R + 2 * I * R + 64;
}

// U (I x R, orthonormal, FP64) <- orthonormal basis of span(S^inner U).
// U32 (I x R), M32 (R x I), Ub16 (I x R) receive the kernel views (may be null).
int dt_als_update(dt_ctx *ctx, const double *S, double *U, int I, int R, int inner, double *ws,
                  int *flag, float *U32, float *M32, uint16_t *Ub16) {
    if (R < 1 || R > ALS_MAXR) return dt_fail(ctx, DT_EUNSUPPORTED, "ALS update: rank %d > %d", R, ALS_MAXR);
    const int nb = (I + ALS_RB - 1) / ALS_RB;
    double *part = ws;
    double *Ri = part + (size_t)nb * R * R;
    double *Yb[2] = {Ri + 2 * R * R, Ri + 2 * R * R + (size_t)I * R};
    const double *cur = U;
    const double *curRi = nullptr;  // Q_0 = U
    for (int k = 0; k < inner; k++) {
        double *nxt = Yb[k & 1];
        DT_TRY(rows(ctx, 0, S, cur, I, R, curRi, nxt, part));
        double *ri = Ri + (size_t)(k & 1) * R * R;
        DT_LAUNCH("als_chol", als_chol_kernel, dim3(1), dim3(ALS_THREADS), 0, part, nb, R, I, 1, ri, flag);
        cur = nxt;
        curRi = ri;
    }
    // Q = Y R^{-1} (orthonormal to ~cond(Y) u), then one plain Cholesky-QR pass
    DT_TRY(rows(ctx, 1, nullptr, cur, I, R, curRi, U, part));
    double *ri = Ri + (size_t)(inner & 1) * R * R;
    DT_LAUNCH("als_chol", als_chol_kernel, dim3(1), dim3(ALS_THREADS), 0, part, nb, R, I, 0, ri, flag);
    return rows(ctx, 1, nullptr, U, I, R, ri, U, nullptr, U32, M32, Ub16);
}
