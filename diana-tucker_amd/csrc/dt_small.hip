// Fused single-workgroup FP64 kernels for the per-mode factor update.
//
// The reference's ALS_ (template/algorithm/tucker/hooi_als.tpp:7-34) is a
// sequence of small BLAS/LAPACK calls (matmulTN, inverse = dgetrf+dgetri,
// matmulNN, reduced_QR = dgeqrf+dorgqr).  Launched one by one with a host
// check after every LAPACK-style call they cost more than the full-tensor
// passes, so here each group of small steps is one launch of one workgroup,
// and failures (singular R x R systems, non-positive Cholesky pivots) raise a
// device flag that the host reads once per sweep:
//   als_prep_kernel   : Gm = LᵀL, Gm^{-1} (in-place Gauss-Jordan, partial
//                       pivoting), P = L Gm^{-1}                (hooi_als.tpp:21-23)
//   als_finish_kernel : GR = Pᵀ W, GR^{-1}, L = W GR^{-1}        (hooi_als.tpp:28-30)
//   cholqr3_kernel    : Q = orth(L) by a shifted Cholesky-QR pass followed by
//                       two plain passes (Fukaya, Kannan, Nakatsukasa, Yamamoto,
//                       Yanagisawa, SISC 2020); positive R diagonal  (hooi_als.tpp:32)
// The I x R panels stream through LDS in row chunks with coalesced loads;
// the R x R matrix lives in LDS (R <= 64).  The W = S P product stays a
// multi-workgroup GEMM.
#include "dt_internal.hpp"

#include <cmath>

#define SMALL_THREADS 512
constexpr int kMaxR = 64;

// Out (R x R, LDS) = Xᵀ Y for X, Y (I x R, column-major, global).  The rows
// are staged CH at a time through LDS (stage: 2 * (CH+1) * R doubles).
__device__ inline double ld_f64(const double *p) { return *p; }
__device__ inline double ld_f64(const float *p) { return (double)*p; }
__device__ inline double ld_f64(const uint16_t *p) { return (double)bf16_to_f32(*p); }

template <typename T>
__device__ void wg_xty(const T *__restrict__ X, const T *__restrict__ Y, int I, int R,
                       double *__restrict__ Out, double *__restrict__ stage, int CH) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int ld = CH + 1;  // odd stride in doubles: column reads hit distinct banks
    double *Xs = stage, *Ys = stage + ld * R;
    // each thread owns up to 8 outputs
    double acc[8];
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] = 0.0;
    const int nout = R * R;
    for (int t0 = 0; t0 < I; t0 += CH) {
        const int ch = (I - t0) < CH ? (I - t0) : CH;
        for (int e = tid; e < CH * R; e += nt) {
            const int r = e % CH, c = e / CH;
            Xs[r + ld * c] = r < ch ? ld_f64(X + (t0 + r) + (size_t)I * c) : 0.0;
            Ys[r + ld * c] = r < ch ? ld_f64(Y + (t0 + r) + (size_t)I * c) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int o = tid + q * nt;
            if (o < nout) {
                const int i = o % R, j = o / R;
                const double *xa = Xs + ld * i, *yb = Ys + ld * j;
                double s = acc[q];
#pragma unroll 8
                for (int r = 0; r < CH; r++) s = fma(xa[r], yb[r], s);
                acc[q] = s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int o = tid + q * nt;
        if (o < nout) Out[o] = acc[q];
    }
    __syncthreads();
}

// Out (I x R, global) = X (I x R, global) M (R x R, LDS).  Rows of X are
// staged CH at a time through LDS before any output of the chunk is written,
// so Out may alias X.
__device__ void wg_xm(const double *X, const double *__restrict__ M, int I, int R, double *Out,
                      double *__restrict__ stage, int CH) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int ld = CH + 1;
    for (int t0 = 0; t0 < I; t0 += CH) {
        const int ch = (I - t0) < CH ? (I - t0) : CH;
        for (int e = tid; e < CH * R; e += nt) {
            const int r = e % CH, c = e / CH;
            stage[r + ld * c] = r < ch ? X[(t0 + r) + (size_t)I * c] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < CH * R; e += nt) {
            const int r = e % CH, j = e / CH;
            if (r < ch) {
                double s = 0.0;
                for (int k = 0; k < R; k++) s = fma(stage[r + ld * k], M[k + R * j], s);
                Out[(t0 + r) + (size_t)I * j] = s;
            }
        }
        __syncthreads();
    }
}

// In-place Gauss-Jordan inverse with partial pivoting of the n x n matrix A
// (LDS, or global memory private to this workgroup); perm holds n ints.
// Returns nonzero on a zero pivot (uniform across the workgroup).
__device__ int wg_inverse_inplace(double *__restrict__ A, int n, int *__restrict__ perm) {
    __shared__ int piv_s, bad_s;
    __shared__ double pinv_s;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) bad_s = 0;
    __syncthreads();
    for (int k = 0; k < n; k++) {
        if (tid < 64) {
            // argmax |A[i][k]| over i >= k in one wave (rows strided by 64);
            // ties keep the lower row
            double v = -1.0;
            int p = n;
            for (int i = k + tid; i < n; i += 64) {
                const double a = fabs(A[i + n * k]);
                if (a > v) {
                    v = a;
                    p = i;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(v, off, 64);
                const int op = __shfl_xor(p, off, 64);
                if (ov > v || (ov == v && op < p)) {
                    v = ov;
                    p = op;
                }
            }
            if (tid == 0) {
                piv_s = p;
                perm[k] = p;
                if (!(v > 0.0) || !isfinite(v)) bad_s = k + 1;
            }
        }
        __syncthreads();
        if (bad_s) return bad_s;
        const int p = piv_s;
        if (p != k)
            for (int j = tid; j < n; j += nt) {
                const double t = A[k + n * j];
                A[k + n * j] = A[p + n * j];
                A[p + n * j] = t;
            }
        __syncthreads();
        if (tid == 0) {
            pinv_s = 1.0 / A[k + n * k];
            A[k + n * k] = 1.0;
        }
        __syncthreads();
        const double pinv = pinv_s;
        for (int j = tid; j < n; j += nt) A[k + n * j] *= pinv;
        __syncthreads();
        // eliminate column k from the other rows (column k holds the multipliers)
        for (int idx = tid; idx < n * n; idx += nt) {
            const int i = idx % n, j = idx / n;
            if (i == k || j == k) continue;
            A[i + n * j] -= A[i + n * k] * A[k + n * j];
        }
        __syncthreads();
        for (int i = tid; i < n; i += nt)
            if (i != k) A[i + n * k] = -A[i + n * k] * A[k + n * k];
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, last first
    for (int k = n - 1; k >= 0; k--) {
        const int p = perm[k];
        if (p != k)
            for (int i = tid; i < n; i += nt) {
                const double t = A[i + n * k];
                A[i + n * k] = A[i + n * p];
                A[i + n * p] = t;
            }
        __syncthreads();
    }
    return 0;
}

// G (n x n SPD) + shift I -> upper Cholesky factor R (in place in G);
// Xo (n x n) <- R^{-1} (upper); invd: n doubles.  Each may be LDS or global
// memory private to this workgroup.  Returns nonzero on a non-positive pivot.
__device__ int wg_chol_rinv(double *__restrict__ G, int n, double shift, double *__restrict__ Xo,
                            double *__restrict__ invd) {
    __shared__ int bad_s;
    __shared__ double rjj_s;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < n; i += nt) G[i + n * i] += shift;
    if (tid == 0) bad_s = 0;
    __syncthreads();
    for (int j = 0; j < n; j++) {
        if (tid == 0) {
            const double d = G[j + n * j];
            if (!(d > 0.0) || !isfinite(d)) {
                bad_s = j + 1;
                rjj_s = 1.0;
            } else {
                rjj_s = sqrt(d);
                G[j + n * j] = rjj_s;
            }
        }
        __syncthreads();
        if (bad_s) return bad_s;
        const double rjj = rjj_s;
        for (int k = j + 1 + tid; k < n; k += nt) G[j + n * k] /= rjj;
        __syncthreads();
        const int mm = n - j - 1;
        for (int idx = tid; idx < mm * mm; idx += nt) {
            const int i = j + 1 + idx % mm, k = j + 1 + idx / mm;
            if (i <= k) G[i + n * k] -= G[j + n * i] * G[j + n * k];
        }
        __syncthreads();
    }
    // Xo = R^{-1}: column j solved by thread j (back substitution), written
    // straight into Xo, which no thread reads during the solve; one division
    // per diagonal entry, shared by all columns
    for (int idx = tid; idx < n * n; idx += nt) Xo[idx] = 0.0;
    for (int i = tid; i < n; i += nt) invd[i] = 1.0 / G[i + n * i];
    __syncthreads();
    for (int j = tid; j < n; j += nt) {
        for (int i = j; i >= 0; i--) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int k = i + 1; k <= j; k++) s -= G[i + n * k] * Xo[k + n * j];
            Xo[i + n * j] = s * invd[i];
        }
    }
    __syncthreads();
    return 0;
}

__device__ inline double *lds_base() {
    extern __shared__ double dyn_lds[];
    return dyn_lds;
}

__device__ __host__ inline int stage_rows(int R) { return R <= 16 ? 64 : (R <= 32 ? 32 : 16); }

// LDS layout: M (R*R) | stage (2*(CH+1)*R); perm in static shared.
__global__ __launch_bounds__(SMALL_THREADS) void als_prep_kernel(const double *__restrict__ L, int I,
                                                                 int R, double *__restrict__ P,
                                                                 int *__restrict__ flag) {
    __shared__ int perm[kMaxR];
    double *M = lds_base(), *stage = M + R * R;
    wg_xty(L, L, I, R, M, stage, stage_rows(R));
    if (wg_inverse_inplace(M, R, perm)) {
        if (threadIdx.x == 0) atomicOr(flag, 1);
        return;
    }
    wg_xm(L, M, I, R, P, stage, stage_rows(R));
}

__global__ __launch_bounds__(SMALL_THREADS) void als_finish_kernel(const double *__restrict__ P,
                                                                   const double *__restrict__ Wm,
                                                                   int I, int R,
                                                                   double *__restrict__ L,
                                                                   int *__restrict__ flag) {
    __shared__ int perm[kMaxR];
    double *M = lds_base(), *stage = M + R * R;
    wg_xty(P, Wm, I, R, M, stage, stage_rows(R));
    if (wg_inverse_inplace(M, R, perm)) {
        if (threadIdx.x == 0) atomicOr(flag, 2);
        return;
    }
    wg_xm(Wm, M, I, R, L, stage, stage_rows(R));
}

// X (m x n) <- orth(X): shifted Cholesky-QR pass + two plain passes.
// LDS: G (n*n) | R^{-1} (n*n) | stage (2*(CH+1)*n).
__global__ __launch_bounds__(SMALL_THREADS) void cholqr3_kernel(double *X, int m, int n,
                                                                int *__restrict__ flag) {
    __shared__ double shift_s;
    __shared__ double invd[kMaxR];
    double *G = lds_base(), *Ri = G + n * n, *stage = Ri + n * n;
    for (int pass = 0; pass < 3; pass++) {
        wg_xty(X, X, m, n, G, stage, stage_rows(n));
        if (threadIdx.x == 0) {
            double shift = 0.0;
            if (pass == 0) {
                double tr = 0.0;
                for (int i = 0; i < n; i++) tr += G[i + n * i];
                shift = 11.0 * ((double)m * n + (double)n * (n + 1)) * 1.1102230246251565e-16 * tr;
            }
            shift_s = shift;
        }
        __syncthreads();
        if (wg_chol_rinv(G, n, shift_s, Ri, invd)) {
            if (threadIdx.x == 0) atomicOr(flag, 4);
            return;
        }
        wg_xm(X, Ri, m, n, X, stage, stage_rows(n));
    }
}

// FP64 MFMA forms of the two products of the fit correction (n <= 64, 8
// waves).  Output tiles of 16 x 16 go to waves round-robin; v_mfma_f64_16x16x4f64
// takes A[i][k] from lane (i, g = k) and B[k][j] from lane (j, g = k), and
// register v of lane (c, g) holds row 4v + g, column c of the tile.  Indices
// past n (and rows past m) read as zero.
__device__ __forceinline__ double ld4_f64(const float *p, int s) { return (double)p[s]; }
__device__ __forceinline__ double ld4_f64(const uint16_t *p, int s) { return (double)bf16_to_f32(p[s]); }
__device__ __forceinline__ double ld4_f64(const double *p, int s) { return p[s]; }

// G (n x n, LDS, column-major) = XᵀX for X (m x n, global, column-major).
// Step s of a 16-row chunk k0 pairs row k = k0 + 4g + s in both operands, so
// each lane reads four consecutive rows of its column.
template <typename T>
__device__ void wg_xtx_mfma(const T *__restrict__ X, int m, int n, double *__restrict__ G) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int c = lane & 15, g = lane >> 4, nt = (n + 15) / 16;
    for (int tile = wv; tile < nt * nt; tile += nw) {
        const int ti = tile % nt, tj = tile / nt;
        const int ca = 16 * ti + c, cb = 16 * tj + c;
        const T *pa = X + (size_t)(ca < n ? ca : n - 1) * m;
        const T *pb = X + (size_t)(cb < n ? cb : n - 1) * m;
        const bool ina = ca < n, inb = cb < n;
        doublex4 acc = {0.0, 0.0, 0.0, 0.0};
        int k0 = 0;
        // four chunks' loads issued before their MFMAs
        for (; k0 + 64 <= m; k0 += 64) {
            double a[4][4], b[4][4];
            if (ina && inb) {
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        a[u][s] = ld4_f64(pa, k0 + 16 * u + 4 * g + s);
                        b[u][s] = ld4_f64(pb, k0 + 16 * u + 4 * g + s);
                    }
            } else {
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        a[u][s] = ina ? ld4_f64(pa, k0 + 16 * u + 4 * g + s) : 0.0;
                        b[u][s] = inb ? ld4_f64(pb, k0 + 16 * u + 4 * g + s) : 0.0;
                    }
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int s = 0; s < 4; s++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][s], b[u][s], acc, 0, 0, 0);
        }
        for (; k0 < m; k0 += 16) {
            double a[4], b[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int k = k0 + 4 * g + s, kc = k < m ? k : m - 1;
                a[s] = (ina && k < m) ? ld4_f64(pa, kc) : 0.0;
                b[s] = (inb && k < m) ? ld4_f64(pb, kc) : 0.0;
            }
#pragma unroll
            for (int s = 0; s < 4; s++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = 16 * ti + 4 * v + g, j = 16 * tj + c;
            if (i < n && j < n) G[i + n * j] = acc[v];
        }
    }
    __syncthreads();
}

// out = x y (n x n, LDS, column-major)
// out = x yᵀ (n x n, LDS, column-major): the B operand is read along a row of
// y (consecutive lanes, consecutive addresses), not down a column, whose
// 512-byte stride at n = 64 puts a column group's 16 lanes on the same banks.
__device__ void wg_small_mm_mfma(const double *__restrict__ x, const double *__restrict__ y, int n,
                                 double *__restrict__ out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int c = lane & 15, g = lane >> 4, nt = (n + 15) / 16;
    for (int tile = wv; tile < nt * nt; tile += nw) {
        const int ti = tile % nt, tj = tile / nt;
        const int i = 16 * ti + c, j = 16 * tj + c;
        doublex4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < n; k0 += 4) {
            const int k = k0 + g;
            const double a = (i < n && k < n) ? x[i + n * k] : 0.0;
            const double b = (k < n && j < n) ? y[j + n * k] : 0.0;  // (yᵀ)[k][j]
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int r = 16 * ti + 4 * v + g;
            if (r < n && j < n) out[r + n * j] = acc[v];
        }
    }
}

// Workgroup q: C_q (n_q x n_q) = (X_qᵀX_q)^{-1}, X_q m_q x n_q column-major in
// FP64, FP32 or bf16 (widened on load).
// out (n x n, LDS) = x y (LDS, n x n column-major); waves over columns, lanes
// over rows, no integer division
__device__ void wg_small_mm(const double *__restrict__ x, const double *__restrict__ y, int n,
                            double *__restrict__ out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int j = wv; j < n; j += nw)
        for (int i = lane; i < n; i += 64) {
            double s0 = 0.0, s1 = 0.0;
            int k = 0;
            for (; k + 1 < n; k += 2) {
                s0 = fma(x[i + n * k], y[k + n * j], s0);
                s1 = fma(x[i + n * (k + 1)], y[k + 1 + n * j], s1);
            }
            if (k < n) s0 = fma(x[i + n * k], y[k + n * j], s0);
            out[i + n * j] = s0 + s1;
        }
}

// Workgroup q: C_q (n_q x n_q) = (X_qᵀX_q)^{-1}, X_q m_q x n_q column-major in
// FP64, FP32 or bf16 (widened on load).  The Gram of a rounded copy of an
// orthonormal factor is I + E with E small: then Newton-Schulz from C = 2I - G,
// C <- C (2I - G C), squares the error each step (5 steps from ||E||_F < 1/4
// reach the FP64 floor) with two small products per step and no pivoting;
// otherwise Gauss-Jordan with partial pivoting.
// LDS: G (n^2) | staging of wg_xty, reused as C, T, T2 (3 n^2).
__global__ __launch_bounds__(SMALL_THREADS) void gram_inverse_batch_kernel(dt_gram_inverse_batch b,
                                                                           int *__restrict__ flag) {
    __shared__ int perm[kMaxR];
    __shared__ double red[SMALL_THREADS / 64];
    const int q = blockIdx.x, m = b.m[q], n = b.n[q];
    const int tid = threadIdx.x, nt = blockDim.x;
    double *G = lds_base(), *stage = G + n * n;
    // 64-row chunks: the staging (2 x 65 x n) fits in the 3 n^2 that Newton-Schulz
    // needs anyway for n >= 44, and the launch provides at least that much
    (void)b.ch;
    if (b.dt[q] == DT_F32)
        wg_xtx_mfma((const float *)b.X[q], m, n, G);
    else if (b.dt[q] == DT_BF16)
        wg_xtx_mfma((const uint16_t *)b.X[q], m, n, G);
    else
        wg_xtx_mfma((const double *)b.X[q], m, n, G);
    // ||G - I||_F^2
    double e2 = 0.0;
    for (int e = tid; e < n * n; e += nt) {
        const double d = G[e] - ((e % n) == (e / n) ? 1.0 : 0.0);
        e2 += d * d;
    }
    for (int o = 32; o > 0; o >>= 1) e2 += __shfl_xor(e2, o);
    if ((tid & 63) == 0) red[tid >> 6] = e2;
    __syncthreads();
    e2 = 0.0;
    for (int w = 0; w < nt / 64; w++) e2 += red[w];
    if (e2 < 1.0 / 16.0) {  // uniform
        double *C = stage, *T = C + n * n, *T2 = T + n * n;
        for (int e = tid; e < n * n; e += nt) C[e] = ((e % n) == (e / n) ? 2.0 : 0.0) - G[e];
        __syncthreads();
        for (int it = 0; it < 5; it++) {
            wg_small_mm_mfma(C, G, n, T);  // T = C Gᵀ = C G (G symmetric)
            __syncthreads();
            wg_small_mm_mfma(C, T, n, T2);  // T2 = C Tᵀ = C G C
            __syncthreads();
            for (int e = tid; e < n * n; e += nt) C[e] = 2.0 * C[e] - T2[e];
            __syncthreads();
        }
        for (int e = tid; e < n * n; e += nt) b.C[q][e] = C[e];
        return;
    }
    if (wg_inverse_inplace(G, n, perm)) {
        if (threadIdx.x == 0) atomicOr(flag, 8);
        return;
    }
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) b.C[q][i] = G[i];
}

// inverse of G (n x n, global) into C (global)
__global__ __launch_bounds__(SMALL_THREADS) void inverse_flag_kernel(const double *__restrict__ G,
                                                                     int n, double *__restrict__ C,
                                                                     int *__restrict__ flag,
                                                                     int bit) {
    __shared__ int perm[kMaxR];
    double *M = lds_base();
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) M[i] = G[i];
    __syncthreads();
    if (wg_inverse_inplace(M, n, perm)) {
        if (threadIdx.x == 0) atomicOr(flag, bit);
        return;
    }
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) C[i] = M[i];
}

// G (n x n SPD, global) -> Rinv (upper inverse Cholesky factor), optional shift.
// LDS: M (n*n) | Xo (n*n).
__global__ __launch_bounds__(SMALL_THREADS) void chol_rinv_kernel(const double *__restrict__ G,
                                                                  int n, int shifted,
                                                                  double m_rows,
                                                                  double *__restrict__ Rinv,
                                                                  int *__restrict__ flag) {
    __shared__ double shift_s;
    __shared__ double invd[kMaxR];
    double *M = lds_base(), *Xo = M + n * n;
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) M[i] = G[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        double shift = 0.0;
        if (shifted) {
            double tr = 0.0;
            for (int i = 0; i < n; i++) tr += M[i + n * i];
            shift = 11.0 * (m_rows * n + (double)n * (n + 1)) * 1.1102230246251565e-16 * tr;
        }
        shift_s = shift;
    }
    __syncthreads();
    if (wg_chol_rinv(M, n, shift_s, Xo, invd)) {
        if (threadIdx.x == 0) atomicOr(flag, 4);
        return;
    }
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) Rinv[i] = Xo[i];
}

// Orders above kMaxR (no LDS copy): G (n x n SPD, global, overwritten by its
// upper Cholesky factor) + the optional shifted-CholeskyQR shift -> Rinv
// (global); invd: n doubles of global scratch.  One workgroup; the matrices
// are private to it, so the barriers order every access.
__global__ __launch_bounds__(SMALL_THREADS) void chol_rinv_global_kernel(double *__restrict__ G, int n,
                                                                         int shifted, double m_rows,
                                                                         double *__restrict__ Rinv,
                                                                         double *__restrict__ invd,
                                                                         int *__restrict__ flag) {
    __shared__ double shift_s;
    if (threadIdx.x == 0) {
        double shift = 0.0;
        if (shifted) {
            double tr = 0.0;
            for (int i = 0; i < n; i++) tr += G[i + (size_t)n * i];
            shift = 11.0 * (m_rows * n + (double)n * (n + 1)) * 1.1102230246251565e-16 * tr;
        }
        shift_s = shift;
    }
    __syncthreads();
    if (wg_chol_rinv(G, n, shift_s, Rinv, invd) && threadIdx.x == 0) atomicOr(flag, 4);
}

// C = G^{-1} (Gauss-Jordan in place in C, global) for orders above kMaxR;
// perm: n ints of global scratch.
__global__ __launch_bounds__(SMALL_THREADS) void inverse_flag_global_kernel(const double *__restrict__ G, int n,
                                                                            double *__restrict__ C,
                                                                            int *__restrict__ perm,
                                                                            int *__restrict__ flag, int bit) {
    for (size_t i = threadIdx.x; i < (size_t)n * n; i += blockDim.x) C[i] = G[i];
    __syncthreads();
    if (wg_inverse_inplace(C, n, perm) && threadIdx.x == 0) atomicOr(flag, bit);
}

// Out (I x R, global) = S (I x I, symmetric, global) X (I x R, global).
// Thread i streams column i of S (= row i, contiguous) with 16-byte loads;
// the rows of X are staged CH at a time through LDS and read as broadcasts.
template <int RMAX>
__device__ void wg_sx(const double *__restrict__ S, const double *__restrict__ X, int I, int R,
                      double *__restrict__ Out, double *__restrict__ stage, int CH) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int ld = CH + 1;
    const bool vec = (I % 2) == 0;
    for (int i0 = 0; i0 < I; i0 += nt) {
        const int i = i0 + tid;
        double acc[RMAX];
#pragma unroll
        for (int j = 0; j < RMAX; j++) acc[j] = 0.0;
        for (int t0 = 0; t0 < I; t0 += CH) {
            const int ch = (I - t0) < CH ? (I - t0) : CH;
            for (int e = tid; e < CH * R; e += nt) {
                const int r = e % CH, c = e / CH;
                stage[r + ld * c] = r < ch ? X[(t0 + r) + (size_t)I * c] : 0.0;
            }
            __syncthreads();
            if (i < I) {
                const double *col = S + (size_t)I * i + t0;
                int r = 0;
                if (vec) {
#pragma unroll 4
                    for (; r + 1 < ch; r += 2) {
                        const double2 v = *(const double2 *)(col + r);
#pragma unroll
                        for (int j = 0; j < RMAX; j++)
                            if (j < R) {
                                acc[j] = fma(v.x, stage[r + ld * j], acc[j]);
                                acc[j] = fma(v.y, stage[r + 1 + ld * j], acc[j]);
                            }
                    }
                }
                for (; r < ch; r++) {
                    const double sv = col[r];
#pragma unroll
                    for (int j = 0; j < RMAX; j++)
                        if (j < R) acc[j] = fma(sv, stage[r + ld * j], acc[j]);
                }
            }
            __syncthreads();
        }
        if (i < I) {
#pragma unroll
            for (int j = 0; j < RMAX; j++)
                if (j < R) Out[i + (size_t)I * j] = acc[j];
        }
    }
    __syncthreads();
}

// The whole per-mode update of ALS_ (hooi_als.tpp:7-34) in one workgroup:
// `inner` steps of P = L (LᵀL)^{-1}, W = S P, L = W (Pᵀ W)^{-1}, then the
// orthonormal basis of span(L) by shifted Cholesky-QR3.  U: in = L_initial
// (orthonormal start), out = Q.  L, P, W: I x R global scratch.
// LDS: M (R*R) | Ri (R*R) | stage (2*(CH+1)*R).
template <int RMAX>
__global__ __launch_bounds__(SMALL_THREADS) void als_fused_kernel(const double *__restrict__ S,
                                                                  double *__restrict__ U, int I,
                                                                  int R, int inner,
                                                                  double *__restrict__ L,
                                                                  double *__restrict__ P,
                                                                  double *__restrict__ W,
                                                                  int *__restrict__ flag) {
    __shared__ int perm[kMaxR];
    __shared__ double shift_s;
    __shared__ double invd[kMaxR];
    const int CH = stage_rows(R);
    double *M = lds_base(), *Ri = M + R * R, *stage = Ri + R * R;
    for (int e = threadIdx.x; e < I * R; e += blockDim.x) L[e] = U[e];
    __syncthreads();
    for (int it = 0; it < inner; it++) {
        const double *Pp = L;
        if (it > 0) {
            wg_xty(L, L, I, R, M, stage, CH);
            if (wg_inverse_inplace(M, R, perm)) {
                if (threadIdx.x == 0) atomicOr(flag, 1);
                return;
            }
            wg_xm(L, M, I, R, P, stage, CH);
            Pp = P;
        }
        wg_sx<RMAX>(S, Pp, I, R, W, stage, CH);
        wg_xty(Pp, W, I, R, M, stage, CH);
        if (wg_inverse_inplace(M, R, perm)) {
            if (threadIdx.x == 0) atomicOr(flag, 2);
            return;
        }
        wg_xm(W, M, I, R, L, stage, CH);
    }
    for (int pass = 0; pass < 3; pass++) {
        wg_xty(L, L, I, R, M, stage, CH);
        if (threadIdx.x == 0) {
            double shift = 0.0;
            if (pass == 0) {
                double tr = 0.0;
                for (int i = 0; i < R; i++) tr += M[i + R * i];
                shift = 11.0 * ((double)I * R + (double)R * (R + 1)) * 1.1102230246251565e-16 * tr;
            }
            shift_s = shift;
        }
        __syncthreads();
        if (wg_chol_rinv(M, R, shift_s, Ri, invd)) {
            if (threadIdx.x == 0) atomicOr(flag, 4);
            return;
        }
        wg_xm(L, Ri, I, R, pass == 2 ? U : L, stage, CH);
    }
}

// ---- host wrappers -----------------------------------------------------------
// One launch for the whole ALS_ + QR when R <= 32 (returns DT_EUNSUPPORTED otherwise).
int dt_als_fused(dt_ctx *ctx, const double *S, double *U, int I, int R, int inner, double *L,
                 double *P, double *W, int *flag) {
    if (R > 32) return DT_EUNSUPPORTED;
    const size_t bytes = (size_t)(2 * R * R + 2 * (stage_rows(R) + 1) * R) * sizeof(double);
    if (R <= 8) {
        DT_LAUNCH("als_fused", als_fused_kernel<8>, dim3(1), dim3(SMALL_THREADS), bytes, S, U, I, R, inner,
                  L, P, W, flag);
    } else if (R <= 16) {
        DT_LAUNCH("als_fused", als_fused_kernel<16>, dim3(1), dim3(SMALL_THREADS), bytes, S, U, I, R, inner,
                  L, P, W, flag);
    } else {
        DT_LAUNCH("als_fused", als_fused_kernel<32>, dim3(1), dim3(SMALL_THREADS), bytes, S, U, I, R, inner,
                  L, P, W, flag);
    }
    return DT_OK;
}

static inline size_t lds_bytes(int R) {
    const int ch = stage_rows(R);
    return (size_t)(R * R + 2 * (ch + 1) * R) * sizeof(double);
}

static int check_r(dt_ctx *ctx, int R) {
    if (R < 1 || R > kMaxR) return dt_fail(ctx, DT_EUNSUPPORTED, "rank %d above the supported %d", R, kMaxR);
    return DT_OK;
}

int dt_als_prep(dt_ctx *ctx, const double *L, int I, int R, double *P, double *work, int *flag) {
    (void)work;
    DT_TRY(check_r(ctx, R));
    DT_LAUNCH("als_small", als_prep_kernel, dim3(1), dim3(SMALL_THREADS), lds_bytes(R), L, I, R, P, flag);
    return DT_OK;
}
int dt_als_finish(dt_ctx *ctx, const double *P, const double *W, int I, int R, double *L,
                  double *work, int *flag) {
    (void)work;
    DT_TRY(check_r(ctx, R));
    DT_LAUNCH("als_small", als_finish_kernel, dim3(1), dim3(SMALL_THREADS), lds_bytes(R), P, W, I, R, L,
              flag);
    return DT_OK;
}
int dt_cholqr3_dev(dt_ctx *ctx, double *X, int m, int n, double *work, int *flag) {
    (void)work;
    DT_TRY(check_r(ctx, n));
    const size_t bytes = (size_t)(2 * n * n + 2 * (stage_rows(n) + 1) * n) * sizeof(double);
    DT_LAUNCH("cholqr_small", cholqr3_kernel, dim3(1), dim3(SMALL_THREADS), bytes, X, m, n, flag);
    return DT_OK;
}
int dt_inverse_flag(dt_ctx *ctx, const double *G, int n, double *C, double *work, int *flag, int bit) {
    if (n > kMaxR) {
        // work: n ints for the pivot rows
        DT_CHECK_ARG(work && C != G, "inverse: order %d needs scratch and C != G", n);
        DT_LAUNCH("als_small", inverse_flag_global_kernel, dim3(1), dim3(SMALL_THREADS), 0, G, n, C,
                  (int *)work, flag, bit);
        return DT_OK;
    }
    DT_TRY(check_r(ctx, n));
    DT_LAUNCH("als_small", inverse_flag_kernel, dim3(1), dim3(SMALL_THREADS), (size_t)n * n * 8, G, n, C,
              flag, bit);
    return DT_OK;
}

int dt_gram_inverse_batch_dev(dt_ctx *ctx, const dt_gram_inverse_batch &b, int *flag) {
    int nmax = 1;
    for (int q = 0; q < b.count; q++) {
        DT_TRY(check_r(ctx, b.n[q]));
        nmax = b.n[q] > nmax ? b.n[q] : nmax;
    }
    if (b.count < 1) return DT_OK;
    // LDS: the Gram and the row staging of wg_xty (64-row chunks when the
    // three Newton-Schulz matrices leave room for them), or those matrices
    dt_gram_inverse_batch bb = b;
    const size_t ns = (size_t)3 * nmax * nmax;
    bb.ch = (size_t)2 * 65 * nmax <= ns ? 64 : stage_rows(nmax);
    const size_t lds = std::max((size_t)nmax * nmax + 2 * (size_t)(bb.ch + 1) * nmax, (size_t)nmax * nmax + ns) *
                       sizeof(double);
    DT_LAUNCH("fit_small", gram_inverse_batch_kernel, dim3((unsigned)b.count), dim3(SMALL_THREADS), lds, bb, flag);
    return DT_OK;
}

int dt_chol_rinv_ws(dt_ctx *ctx, double *G, int n, bool shifted, double m_rows, double *Rinv, double *invd,
                    int *flag) {
    if (n <= kMaxR) return dt_chol_rinv_dev(ctx, G, n, shifted, m_rows, Rinv, flag);
    DT_LAUNCH("cholqr_small", chol_rinv_global_kernel, dim3(1), dim3(SMALL_THREADS), 0, G, n, shifted ? 1 : 0,
              m_rows, Rinv, invd, flag);
    return DT_OK;
}

int dt_chol_rinv_dev(dt_ctx *ctx, const double *G, int n, bool shifted, double m_rows, double *Rinv, int *flag) {
    DT_TRY(check_r(ctx, n));
    DT_LAUNCH("cholqr_small", chol_rinv_kernel, dim3(1), dim3(SMALL_THREADS), 2 * n * n * 8, G, n, shifted ? 1 : 0,
              m_rows, Rinv, flag);
    return DT_OK;
}

// Device-only orthonormalisation of X (m x n): shifted + 2 plain Cholesky-QR
// passes.  One workgroup when m n^2 is small, otherwise the Gram and the
// X R^{-1} products run as multi-workgroup GEMMs.  work: >= 2 n*n + m*n doubles.

int dt_orth_dev(dt_ctx *ctx, double *X, size_t m, size_t n, double *work, int *flag) {
    if ((double)m * n * n <= (double)(1 << 21) && n <= 32)
        return dt_cholqr3_dev(ctx, X, (int)m, (int)n, work, flag);
    double *G = work, *Rinv = work + n * n, *T = work + 2 * n * n;
    for (int pass = 0; pass < 3; pass++) {
        DT_TRY(dt_gemm_f64_internal(ctx, 1, 0, n, n, m, 1.0, X, m, X, m, 0.0, G, n));
        // T (m n >= n doubles) is free until the product below: the reciprocal
        // diagonal of the global-memory factorisation goes there
        DT_TRY(dt_chol_rinv_ws(ctx, G, (int)n, pass == 0, (double)m, Rinv, T, flag));
        DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, m, n, n, 1.0, X, m, Rinv, n, 0.0, T, m));
        DT_HIP(hipMemcpyAsync(X, T, m * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    return DT_OK;
}
