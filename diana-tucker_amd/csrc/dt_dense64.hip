// FP64 small dense linear algebra on the GPU (no vendor library).
//
// Reference counterparts (src/operator/operator_double.cpp):
//   matmulNN/NT/TN :157-216  (cblas_dgemm)     -> gemm_f64_kernel (+ split-K)
//   inverse        :84-103   (dgetrf+dgetri)   -> gj_inverse_kernel (Gauss-Jordan,
//                                                 partial pivoting, one workgroup)
//   QR             :131-155  (dgeqrf+dorgqr)   -> Cholesky-QR2 with a shifted
//                                                 first pass when the Gram is not
//                                                 numerically positive definite
//                                                 (Fukaya et al., SISC 2020)
//   orth1/orth2    :218-277  (MGS)             -> the same QR on the (transposed)
//                                                 matrix: equal span, positive R diag
// plus a cyclic-Jacobi symmetric eigensolver (Golub & Van Loan, 8.5) for
// Rayleigh-Ritz blocks and an oversampled subspace iteration for the leading
// eigenvectors of larger Gram matrices (HOSVD start factors).
#include "dt_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

static inline dim3 gs_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 262144) g = 262144;
    if (g < 1) g = 1;
    return dim3((unsigned)g);
}

// ---------------------------------------------------------------------------
// GEMM: C = alpha op(A) op(B) + beta C (column-major) on FP64 MFMA.  A
// workgroup computes a 64 x 64 tile of C over its k range (blockIdx.z); wave w
// owns the 32 x 32 quarter (w & 1, w >> 1) as 2 x 2 v_mfma_f64_16x16x4f64
// tiles.  k is staged 16 at a time through LDS, op(A) as [row][k] and op(B) as
// [k][column], so a transpose only changes which threads load which entries
// (consecutive threads always read consecutive addresses).  MFMA step s, lane
// (c, g): op(A)[row c][k 4s + g] and op(B)[k 4s + g][column c]; register v of
// the result is row 4v + g, column c.  With part != nullptr the workgroup
// writes its unscaled partial sum over its k range into part[z][i + m*j].
__global__ __launch_bounds__(256) void gemm_f64_kernel(int tA, int tB, int64_t m, int64_t n,
                                                       int64_t k, double alpha,
                                                       const double *__restrict__ A, int64_t lda,
                                                       const double *__restrict__ B, int64_t ldb,
                                                       double beta, double *__restrict__ C,
                                                       int64_t ldc, int64_t kchunk,
                                                       double *__restrict__ part) {
    __shared__ double As[64][17];
    __shared__ double Bs[16][65];
    const int64_t i0 = (int64_t)blockIdx.x * 64;
    const int64_t j0 = (int64_t)blockIdx.y * 64;
    const int64_t ks = (int64_t)blockIdx.z * kchunk;
    const int64_t ke = ks + kchunk < k ? ks + kchunk : k;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
    doublex4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};
    for (int64_t k0 = ks; k0 < ke; k0 += 16) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 256 * q;
            const int i = tA ? e >> 4 : e & 63, kk = tA ? e & 15 : e >> 6;
            const int64_t gi = i0 + i, gk = k0 + kk;
            double v = 0.0;
            if (gi < m && gk < ke) v = tA ? A[gk + lda * gi] : A[gi + lda * gk];
            As[i][kk] = v;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 256 * q;
            const int j = tB ? e & 63 : e >> 4, kk = tB ? e >> 6 : e & 15;
            const int64_t gj = j0 + j, gk = k0 + kk;
            double v = 0.0;
            if (gj < n && gk < ke) v = tB ? B[gj + ldb * gk] : B[gk + ldb * gj];
            Bs[kk][j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const double a0 = As[wm + c][4 * s + g], a1 = As[wm + 16 + c][4 * s + g];
            const double b0 = Bs[4 * s + g][wn + c], b1 = Bs[4 * s + g][wn + 16 + c];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t i = i0 + wm + 16 * a + 4 * v + g, j = j0 + wn + 16 * b + c;
                if (i >= m || j >= n) continue;
                if (part) {
                    part[(int64_t)blockIdx.z * m * n + i + m * j] = acc[a][b][v];
                } else {
                    const double old = beta == 0.0 ? 0.0 : beta * C[i + ldc * j];
                    C[i + ldc * j] = alpha * acc[a][b][v] + old;
                }
            }
}

__global__ void gemm_reduce(const double *__restrict__ part, int nsplit, int64_t m, int64_t n,
                            double alpha, double beta, double *__restrict__ C, int64_t ldc) {
    const int64_t total = m * n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        // four independent sums so that four loads are in flight
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int q = 0;
        for (; q + 3 < nsplit; q += 4) {
            a0 += part[(int64_t)q * total + idx];
            a1 += part[(int64_t)(q + 1) * total + idx];
            a2 += part[(int64_t)(q + 2) * total + idx];
            a3 += part[(int64_t)(q + 3) * total + idx];
        }
        for (; q < nsplit; q++) a0 += part[(int64_t)q * total + idx];
        const double s = (a0 + a1) + (a2 + a3);
        const int64_t i = idx % m, j = idx / m;
        const double old = beta == 0.0 ? 0.0 : beta * C[i + ldc * j];
        C[i + ldc * j] = alpha * s + old;
    }
}

int dt_gemm_f64_internal(dt_ctx *ctx, int tA, int tB, size_t m, size_t n, size_t k, double alpha,
                         const double *A, size_t lda, const double *B, size_t ldb, double beta,
                         double *C, size_t ldc) {
    if (m == 0 || n == 0) return DT_OK;
    // split k so that about 512 workgroups run, each over at least 64 of k
    const int64_t tiles = ceil_div((int64_t)m, 64) * ceil_div((int64_t)n, 64);
    int64_t nsplit = 1;
    if (tiles < 512 && k >= 128)
        nsplit = std::min<int64_t>(ceil_div((int64_t)k, 64), ceil_div((int64_t)512, tiles));
    int64_t kchunk = ceil_div(ceil_div((int64_t)k, nsplit), 16) * 16;
    if (kchunk == 0) kchunk = 16;
    nsplit = std::max<int64_t>(1, ceil_div((int64_t)k, kchunk));
    dim3 grid((unsigned)ceil_div((int64_t)m, 64), (unsigned)ceil_div((int64_t)n, 64),
              (unsigned)nsplit);
    if (nsplit == 1) {
        DT_LAUNCH("gemm_f64", gemm_f64_kernel, grid, dim3(256), 0, tA, tB, (int64_t)m, (int64_t)n,
                  (int64_t)k, alpha, A, (int64_t)lda, B, (int64_t)ldb, beta, C, (int64_t)ldc, kchunk,
                  (double *)nullptr);
        return DT_OK;
    }
    double *part = (double *)dt_ws(ctx, WS_GEMM_PART, (size_t)(nsplit * m * n) * sizeof(double));
    if (!part) return DT_ENOMEM;
    DT_LAUNCH("gemm_f64", gemm_f64_kernel, grid, dim3(256), 0, tA, tB, (int64_t)m, (int64_t)n,
              (int64_t)k, alpha, A, (int64_t)lda, B, (int64_t)ldb, beta, C, (int64_t)ldc, kchunk,
              part);
    DT_LAUNCH("gemm_reduce", gemm_reduce, gs_grid((int64_t)(m * n)), dim3(256), 0, part, (int)nsplit,
              (int64_t)m, (int64_t)n, alpha, beta, C, (int64_t)ldc);
    return DT_OK;
}

extern "C" int dt_gemm_f64(dt_ctx *ctx, int transA, int transB, size_t m, size_t n, size_t k,
                           double alpha, const double *A, size_t lda, const double *B, size_t ldb,
                           double beta, double *C, size_t ldc) {
    DT_CHECK_ARG(ctx && A && B && C, "dt_gemm_f64: null argument");
    DT_CHECK_ARG(ldc >= m && lda >= (transA ? k : m) && ldb >= (transB ? n : k),
                 "dt_gemm_f64: leading dimension too small");
    return dt_gemm_f64_internal(ctx, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// ---------------------------------------------------------------------------
// Cholesky G = R^T R (R upper) and R^{-1}; one workgroup, global memory.
// shift is added to the diagonal first.  info[0] = 0 or the failing pivot+1.
__global__ __launch_bounds__(256) void chol_upper_kernel(const double *__restrict__ G, int n,
                                                         double shift, double *__restrict__ R,
                                                         double *__restrict__ Rinv,
                                                         int *__restrict__ info) {
    __shared__ int bad;
    __shared__ double rjj_s;
    const int tid = threadIdx.x, nt = blockDim.x;
    // W = upper(G) + shift I, stored in R.
    for (int idx = tid; idx < n * n; idx += nt) {
        const int i = idx % n, j = idx / n;
        R[idx] = (i <= j) ? G[idx] + (i == j ? shift : 0.0) : 0.0;
    }
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < n; j++) {
        if (tid == 0) {
            const double d = R[j + n * j];
            if (!(d > 0.0) || !isfinite(d)) {
                bad = j + 1;
                rjj_s = 1.0;
            } else {
                rjj_s = sqrt(d);
                R[j + n * j] = rjj_s;
            }
        }
        __syncthreads();
        if (bad) break;
        const double rjj = rjj_s;
        for (int k = j + 1 + tid; k < n; k += nt) R[j + n * k] /= rjj;
        __syncthreads();
        // trailing update of the upper triangle: W[i][k] -= R[j][i] R[j][k], j < i <= k
        const int m = n - j - 1;
        for (int idx = tid; idx < m * m; idx += nt) {
            const int i = j + 1 + idx % m, k = j + 1 + idx / m;
            if (i <= k) R[i + n * k] -= R[j + n * i] * R[j + n * k];
        }
        __syncthreads();
    }
    if (tid == 0) info[0] = bad;
    if (bad) return;
    // Rinv: solve R X = I column by column (upper triangular back substitution).
    for (int c = tid; c < n; c += nt) {
        for (int i = 0; i < n; i++) Rinv[i + n * c] = 0.0;
        for (int i = c; i >= 0; i--) {
            double s = (i == c) ? 1.0 : 0.0;
            for (int k = i + 1; k <= c; k++) s -= R[i + n * k] * Rinv[k + n * c];
            Rinv[i + n * c] = s / R[i + n * i];
        }
    }
}

// The same factorisation with the matrix held in LDS (n <= LDS_NMAX): one
// workgroup, two barriers per column instead of global-memory round trips.
constexpr int LDS_NMAX = 96;
__global__ __launch_bounds__(512) void chol_lds_kernel(const double *__restrict__ G, int n, double shift,
                                                      double *__restrict__ R, double *__restrict__ Rinv,
                                                      int *__restrict__ info) {
    // W (upper part: the working matrix, ending as R) and Z (lower part: the
    // identity under the same row operations, ending as R^{-T}) with an odd
    // leading dimension ld, so threads that walk a row across columns (stride
    // ld doubles) fall on different LDS banks.  Pivot j: every row i > j of
    // the trailing block loses (W[j][i] / W[j][j]) times row j, on W and on Z,
    // then row j is scaled by W[j][j]^{-1/2}; so E G = R for the product E of
    // these operations, E = R^{-T} is what Z becomes, and no triangular solve
    // is needed.  Threads form a 32 x 16 grid (no index divisions).
    __shared__ double W[LDS_NMAX * (LDS_NMAX + 1)];
    __shared__ double Z[LDS_NMAX * (LDS_NMAX + 1)];
    __shared__ int bad;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int tx = tid & 31, ty = tid >> 5;  // 32 x 16
    const int ld = n | 1;
    for (int j = ty; j < n; j += 16)
        for (int i = tx; i < n; i += 32) {
            W[i + ld * j] = (i <= j) ? G[i + (size_t)n * j] + (i == j ? shift : 0.0) : 0.0;
            Z[i + ld * j] = i == j ? 1.0 : 0.0;
        }
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < n; j++) {
        const double d = W[j + ld * j];
        const bool ok = d > 0.0 && isfinite(d);
        if (!ok) {
            if (tid == 0) bad = j + 1;
            break;  // uniform: every thread read the same d
        }
        const double inv_d = 1.0 / d;
        // trailing block of W (upper part, j < i <= k) and rows i > j of Z
        // (columns 0..j, Z[i][c] at Z[i + ld c])
        for (int k = j + 1 + ty; k < n; k += 16) {
            const double fk = W[j + ld * k] * inv_d;
            for (int i = j + 1 + tx; i <= k; i += 32) W[i + ld * k] -= W[j + ld * i] * fk;
        }
        for (int c = ty; c <= j; c += 16) {
            const double zj = Z[j + ld * c];
            for (int i = j + 1 + tx; i < n; i += 32) Z[i + ld * c] -= W[j + ld * i] * inv_d * zj;
        }
        __syncthreads();
        const double rs = 1.0 / sqrt(d);
        for (int k = j + 1 + tid; k < n; k += nt) W[j + ld * k] *= rs;
        for (int c = tid; c <= j; c += nt) Z[j + ld * c] *= rs;
        if (tid == 0) W[j + ld * j] = d * rs;
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0) info[0] = bad;
    if (bad) return;
    // R = upper part of W; R^{-1}[i][c] = Z[c][i] (upper)
    for (int j = ty; j < n; j += 16)
        for (int i = tx; i < n; i += 32) {
            R[i + (size_t)n * j] = i <= j ? W[i + ld * j] : 0.0;
            Rinv[i + (size_t)n * j] = i <= j ? Z[j + ld * i] : 0.0;
        }
}

static int launch_chol(dt_ctx *ctx, const double *G, size_t n, double shift, double *R, double *Rinv,
                       int *dinfo) {
    if (n <= (size_t)LDS_NMAX) {
        DT_LAUNCH("chol", chol_lds_kernel, dim3(1), dim3(512), 0, G, (int)n, shift, R, Rinv, dinfo);
    } else {
        DT_LAUNCH("chol", chol_upper_kernel, dim3(1), dim3(256), 0, G, (int)n, shift, R, Rinv, dinfo);
    }
    return DT_OK;
}

int dt_chol_upper(dt_ctx *ctx, const double *G, size_t n, double *R, double *Rinv, int *info_host) {
    int *dinfo = (int *)dt_ws(ctx, WS_STATUS, 64);
    if (!dinfo) return DT_ENOMEM;
    DT_TRY(launch_chol(ctx, G, n, 0.0, R, Rinv, dinfo));
    int *pin = (int *)ctx->pinned;
    DT_HIP(hipMemcpyAsync(pin, dinfo, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    *info_host = pin[0];
    return DT_OK;
}

// trace(G) -> out[0] (single thread; G is small)
__global__ void trace_kernel(const double *__restrict__ G, int n, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += fabs(G[i + n * i]);
        out[0] = s;
    }
}

// C = A B for small square upper-triangular n x n (R accumulation), general.
__global__ void small_matmul(const double *__restrict__ A, const double *__restrict__ B,
                             double *__restrict__ C, int n) {
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * n; idx += gridDim.x * blockDim.x) {
        const int i = idx % n, j = idx / n;
        double s = 0.0;
        for (int k = 0; k < n; k++) s += A[i + n * k] * B[k + n * j];
        C[idx] = s;
    }
}

__global__ void set_identity(double *__restrict__ A, int n) {
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * n; idx += gridDim.x * blockDim.x)
        A[idx] = (idx % n == idx / n) ? 1.0 : 0.0;
}

// ---- tall-skinny pieces of Cholesky-QR (n <= TS_NMAX, m large) ---------------
// Partial Grams on v_mfma_f64_16x16x4f64: each wave streams a contiguous run of
// rows straight from global memory, 16 rows per iteration, lane (c, g) loading
// Q[r + 4g + j, c] (and Q[r + 4g + j, 16 + c] when n > 16) for j = 0..3 and
// feeding it as both operands of step j, so D accumulates the 16 x 16 Gram
// tiles.  The four waves' tiles are added in wave order and written to
// part[block][i + n j].
constexpr int TS_NMAX = 32;
typedef double doublex4 __attribute__((ext_vector_type(4)));
template <bool HI>
__global__ __launch_bounds__(256) void ts_gram_kernel(const double *__restrict__ Q, int64_t m, int n,
                                                      int64_t rows_per_wave, double *__restrict__ part) {
    constexpr int NT = HI ? 3 : 1;  // tiles (lo,lo), (lo,hi), (hi,hi)
    __shared__ double red[4][NT][4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int64_t w = (int64_t)blockIdx.x * 4 + wv;
    const int64_t r0 = w * rows_per_wave;
    const int64_t r1 = r0 + rows_per_wave < m ? r0 + rows_per_wave : m;
    const bool clo = c < n, chi = HI && c + 16 < n;
    const double *__restrict__ qlo = Q + m * (clo ? c : 0);
    const double *__restrict__ qhi = Q + m * (chi ? c + 16 : 0);
    doublex4 dll = {0.0, 0.0, 0.0, 0.0}, dlh = {0.0, 0.0, 0.0, 0.0}, dhh = {0.0, 0.0, 0.0, 0.0};
    // 64 rows per iteration: every load is issued (address clamped into the
    // matrix) before any is used, and rows past r1 or columns past n read as 0
    for (int64_t r = r0; r < r1; r += 64) {
        double lo[16], hi[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int64_t rr = r + 16 * (j >> 2) + 4 * g + (j & 3);
            const int64_t rc = rr < m ? rr : m - 1;
            lo[j] = qlo[rc];
            if (HI) hi[j] = qhi[rc];
        }
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int64_t rr = r + 16 * (j >> 2) + 4 * g + (j & 3);
            const double a = (clo && rr < r1) ? lo[j] : 0.0;
            dll = __builtin_amdgcn_mfma_f64_16x16x4f64(a, a, dll, 0, 0, 0);
            if (HI) {
                const double b = (chi && rr < r1) ? hi[j] : 0.0;
                dlh = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, dlh, 0, 0, 0);
                dhh = __builtin_amdgcn_mfma_f64_16x16x4f64(b, b, dhh, 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        red[wv][0][v][lane] = dll[v];
        if (HI) {
            red[wv][1][v][lane] = dlh[v];
            red[wv][2][v][lane] = dhh[v];
        }
    }
    __syncthreads();
    // thread (v, lane): register v of lane (c, g) holds D[4v + g][c] (the FP64
    // 16x16x4 output layout, unlike the FP32 one's D[4g + v][c])
    const int v = threadIdx.x >> 6;
    double *out = part + (int64_t)blockIdx.x * n * n;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const double sum = ((red[0][t][v][lane] + red[1][t][v][lane]) + red[2][t][v][lane]) + red[3][t][v][lane];
        const int i = 4 * v + g + (t == 2 ? 16 : 0), j = c + (t >= 1 ? 16 : 0);
        if (i < n && j < n) {
            out[i + n * j] = sum;
            if (t == 1) out[j + n * i] = sum;
        }
    }
}

// out[e] = sum over p of part[p][e], e < ne: 16 entries per workgroup, 16
// groups of threads each summing every 16th partial, combined in group order.
__global__ __launch_bounds__(256) void part_reduce_kernel(const double *__restrict__ part, int nparts, int ne,
                                                          double *__restrict__ out) {
    __shared__ double red[16][17];
    const int c = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int e = blockIdx.x * 16 + c;
    // eight independent sums per thread so that eight loads are in flight
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (e < ne) {
        int p = grp;
        for (; p + 7 * 16 < nparts; p += 8 * 16)
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] += part[(int64_t)(p + 16 * u) * ne + e];
        for (; p < nparts; p += 16) a[0] += part[(int64_t)p * ne + e];
    }
    red[grp][c] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    __syncthreads();
    if (grp == 0 && e < ne) {
        double t = 0.0;
        for (int g = 0; g < 16; g++) t += red[g][c];
        out[e] = t;
    }
}

// Q[r, :] <- Q[r, :] Rinv in place (Rinv upper triangular n x n, column-major),
// one row per thread.
template <int NMAX>
__global__ __launch_bounds__(256) void ts_rmul_kernel(double *__restrict__ Q, int64_t m, int n,
                                                      const double *__restrict__ Rinv) {
    __shared__ double Rs[NMAX][NMAX + 1];  // Rs[k][j] = Rinv[k + n j]
    for (int e = threadIdx.x; e < NMAX * NMAX; e += 256) {
        const int k = e % NMAX, j = e / NMAX;
        Rs[k][j] = (k < n && j < n && k <= j) ? Rinv[k + n * j] : 0.0;
    }
    __syncthreads();
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < m; r += (int64_t)gridDim.x * 256) {
        double q[NMAX];
#pragma unroll
        for (int c = 0; c < NMAX; c++) q[c] = c < n ? Q[r + m * c] : 0.0;
#pragma unroll
        for (int j = 0; j < NMAX; j++) {
            if (j < n) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k <= j; k++) s = fma(q[k], Rs[k][j], s);
                Q[r + m * j] = s;
            }
        }
    }
}

// G = QᵀQ (n x n) for a tall-skinny column-major Q, FP64, fixed summation order.
static int ts_gram(dt_ctx *ctx, const double *Q, size_t m, size_t n, double *G) {
    // runs of 64-row chunks, at most 4096 waves
    const int64_t chunks = ceil_div((int64_t)m, 64);
    const int64_t nwaves = std::max<int64_t>(1, std::min<int64_t>(4096, ceil_div(chunks, 4)));
    const int64_t rpw = ceil_div(chunks, nwaves) * 64;
    const int64_t ng = ceil_div(ceil_div((int64_t)m, rpw), 4);
    double *part = (double *)dt_ws(ctx, WS_GEMM_PART, (size_t)(ng * n * n) * sizeof(double));
    if (!part) return DT_ENOMEM;
    if (n <= 16) {
        DT_LAUNCH("ts_gram", ts_gram_kernel<false>, dim3((unsigned)ng), dim3(256), 0, Q, (int64_t)m, (int)n, rpw, part);
    } else {
        DT_LAUNCH("ts_gram", ts_gram_kernel<true>, dim3((unsigned)ng), dim3(256), 0, Q, (int64_t)m, (int)n, rpw, part);
    }
    DT_LAUNCH("ts_gram", part_reduce_kernel, dim3((unsigned)ceil_div((int64_t)(n * n), 16)), dim3(256), 0,
              (const double *)part, (int)ng, (int)(n * n), G);
    return DT_OK;
}

int dt_part_reduce(dt_ctx *ctx, const double *part, int nparts, int ne, double *out) {
    DT_LAUNCH("ts_gram", part_reduce_kernel, dim3((unsigned)ceil_div((int64_t)ne, 16)), dim3(256), 0, part, nparts, ne,
              out);
    return DT_OK;
}

static int ts_rmul(dt_ctx *ctx, double *Q, size_t m, size_t n, const double *Rinv) {
    const dim3 grid((unsigned)std::min<int64_t>(ceil_div((int64_t)m, 256), 8192));
    if (n <= 16) {
        DT_LAUNCH("ts_rmul", ts_rmul_kernel<16>, grid, dim3(256), 0, Q, (int64_t)m, (int)n, Rinv);
    } else {
        DT_LAUNCH("ts_rmul", ts_rmul_kernel<32>, grid, dim3(256), 0, Q, (int64_t)m, (int)n, Rinv);
    }
    return DT_OK;
}

// Reduced QR through Cholesky-QR2 (a shifted Cholesky-QR first pass when the
// plain one breaks down).  Q may alias A.
extern "C" int dt_qr_f64(dt_ctx *ctx, double *Q, double *R, const double *A, size_t m, size_t n) {
    DT_CHECK_ARG(ctx && Q && A, "dt_qr_f64: null argument");
    DT_CHECK_ARG(m >= n && n >= 1, "dt_qr_f64: need m >= n >= 1 (got %zu x %zu)", m, n);
    const size_t nn = n * n;
    // Q only, 16 < n <= 64: shifted Cholesky-QR3 over many workgroups with one
    // host wait for its breakdown flag (the GEMM path below waits after every
    // Cholesky and runs a tiled GEMM with few tiles at these shapes)
    if (!R && n > 16 && n <= 64 && m >= 256 && m < ((size_t)1 << 31) && !(n <= (size_t)TS_NMAX && m >= 4096)) {
        const size_t wsd = dt_cholqr3_ws_doubles(m, n) + 8;
        double *ws = (double *)dt_ws(ctx, WS_CHOLQR, (wsd + 8) * sizeof(double));
        if (!ws) return DT_ENOMEM;
        int *dflag = (int *)(ws + wsd);
        if (Q != A) DT_HIP(hipMemcpyAsync(Q, A, m * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        DT_TRY(dt_zero(ctx, dflag, sizeof(int)));
        DT_TRY(dt_cholqr3_rows(ctx, Q, (int)m, (int)n, ws, dflag, true));
        int *pin = (int *)ctx->pinned;
        DT_HIP(hipMemcpyAsync(pin, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        DT_HIP(hipStreamSynchronize(ctx->stream));
        if (pin[0] != 0) return dt_fail(ctx, DT_ENUMERIC, "dt_qr_f64: rank-deficient or non-finite input");
        return DT_OK;
    }
    double *G = (double *)dt_ws(ctx, WS_SMALL0, nn * sizeof(double));
    double *Rk = (double *)dt_ws(ctx, WS_SMALL1, nn * sizeof(double));
    double *Rki = (double *)dt_ws(ctx, WS_SMALL2, nn * sizeof(double));
    // tall-skinny: streaming Gram and in-place row update instead of the tiled GEMM
    const bool ts = n <= (size_t)TS_NMAX && m >= 4096;
    double *Racc = (double *)dt_ws(ctx, WS_QR_T1, 2 * nn * sizeof(double) + (ts ? 0 : m * n * sizeof(double)));
    if (!G || !Rk || !Rki || !Racc) return DT_ENOMEM;
    double *Rtmp = Racc + nn;
    double *T = Rtmp + nn;  // m x n temporary
    if (Q != A) DT_HIP(hipMemcpyAsync(Q, A, m * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if (R) DT_LAUNCH("qr_small", set_identity, dim3(1), dim3(256), 0, Racc, (int)n);
    int passes = 2;
    bool shifted = false;
    for (int pass = 0; pass < passes; pass++) {
        if (ts) {
            DT_TRY(ts_gram(ctx, Q, m, n, G));
        } else {
            DT_TRY(dt_gemm_f64_internal(ctx, 1, 0, n, n, m, 1.0, Q, m, Q, m, 0.0, G, n));
        }
        int info = 0;
        DT_TRY(dt_chol_upper(ctx, G, n, Rk, Rki, &info));
        if (info != 0) {
            if (shifted) return dt_fail(ctx, DT_ENUMERIC, "dt_qr_f64: Cholesky breakdown after shift");
            // shift = 11 (m n + n (n+1)) u ||A||_2^2, bounded above by trace(G)
            double *dtr = (double *)dt_ws(ctx, WS_STATUS, 64);
            DT_LAUNCH("qr_small", trace_kernel, dim3(1), dim3(64), 0, G, (int)n, dtr);
            DT_HIP(hipMemcpyAsync(ctx->pinned, dtr, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            DT_HIP(hipStreamSynchronize(ctx->stream));
            const double tr = ctx->pinned[0];
            if (!(tr > 0.0) || !std::isfinite(tr))
                return dt_fail(ctx, DT_ENUMERIC, "dt_qr_f64: zero or non-finite matrix");
            const double shift = 11.0 * ((double)m * n + (double)n * (n + 1)) * 1.1102230246251565e-16 * tr;
            int *dinfo = (int *)dt_ws(ctx, WS_STATUS, 64);
            DT_TRY(launch_chol(ctx, G, n, shift, Rk, Rki, dinfo));
            int *pin = (int *)ctx->pinned;
            DT_HIP(hipMemcpyAsync(pin, dinfo, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            DT_HIP(hipStreamSynchronize(ctx->stream));
            if (pin[0] != 0)
                return dt_fail(ctx, DT_ENUMERIC, "dt_qr_f64: rank-deficient input (pivot %d)", pin[0]);
            shifted = true;
            passes = 3;
        }
        if (ts) {
            DT_TRY(ts_rmul(ctx, Q, m, n, Rki));
        } else {
            DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, m, n, n, 1.0, Q, m, Rki, n, 0.0, T, m));
            DT_HIP(hipMemcpyAsync(Q, T, m * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        }
        if (R) {
            DT_LAUNCH("qr_small", small_matmul, dim3(16), dim3(256), 0, Rk, Racc, Rtmp, (int)n);
            DT_HIP(hipMemcpyAsync(Racc, Rtmp, nn * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        }
    }
    if (R) DT_HIP(hipMemcpyAsync(R, Racc, nn * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return DT_OK;
}

extern "C" int dt_orth_rows_f64(dt_ctx *ctx, double *A, size_t len1, size_t len2) {
    DT_CHECK_ARG(ctx && A, "dt_orth_rows_f64: null argument");
    DT_CHECK_ARG(len2 >= len1, "dt_orth_rows_f64: need len2 >= len1");
    double *T = (double *)dt_ws(ctx, WS_QR_T0, len1 * len2 * sizeof(double));
    if (!T) return DT_ENOMEM;
    DT_TRY(dt_transpose_f64(ctx, T, A, len1, len2));  // T: len2 x len1
    DT_TRY(dt_qr_f64(ctx, T, nullptr, T, len2, len1));
    return dt_transpose_f64(ctx, A, T, len2, len1);
}

// ---------------------------------------------------------------------------
// Gauss-Jordan inverse with partial pivoting, one workgroup.  W: n x 2n scratch.
__global__ __launch_bounds__(256) void gj_inverse_kernel(const double *__restrict__ A, int n,
                                                         double *__restrict__ W,
                                                         double *__restrict__ C,
                                                         int *__restrict__ info) {
    __shared__ int piv;
    __shared__ int bad;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int ld = n;
    for (int idx = tid; idx < 2 * n * n; idx += nt) {
        const int i = idx % n, j = idx / n;
        W[i + ld * j] = (j < n) ? A[i + n * j] : ((j - n) == i ? 1.0 : 0.0);
    }
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int col = 0; col < n; col++) {
        if (tid == 0) {
            int p = col;
            double best = fabs(W[col + ld * col]);
            for (int i = col + 1; i < n; i++) {
                const double v = fabs(W[i + ld * col]);
                if (v > best) {
                    best = v;
                    p = i;
                }
            }
            piv = p;
            if (!(best > 0.0) || !isfinite(best)) bad = col + 1;
        }
        __syncthreads();
        if (bad) break;
        const int p = piv;
        if (p != col) {
            for (int j = tid; j < 2 * n; j += nt) {
                const double t = W[col + ld * j];
                W[col + ld * j] = W[p + ld * j];
                W[p + ld * j] = t;
            }
        }
        __syncthreads();
        const double d = W[col + ld * col];
        __syncthreads();
        for (int j = tid; j < 2 * n; j += nt) W[col + ld * j] /= d;
        __syncthreads();
        // eliminate column col from every other row
        for (int idx = tid; idx < n * 2 * n; idx += nt) {
            const int i = idx % n, j = idx / n;
            if (i == col || j == col) continue;
            W[i + ld * j] -= W[i + ld * col] * W[col + ld * j];
        }
        __syncthreads();
        for (int i = tid; i < n; i += nt)
            if (i != col) W[i + ld * col] = 0.0;
        __syncthreads();
    }
    if (tid == 0) info[0] = bad;
    if (bad) return;
    for (int idx = tid; idx < n * n; idx += nt) {
        const int i = idx % n, j = idx / n;
        C[idx] = W[i + ld * (n + j)];
    }
}

extern "C" int dt_inverse_f64(dt_ctx *ctx, double *C, const double *A, size_t n) {
    DT_CHECK_ARG(ctx && C && A && n >= 1, "dt_inverse_f64: bad argument");
    double *W = (double *)dt_ws(ctx, WS_SMALL3, 2 * n * n * sizeof(double));
    int *dinfo = (int *)dt_ws(ctx, WS_STATUS, 64);
    if (!W || !dinfo) return DT_ENOMEM;
    DT_LAUNCH("inverse", gj_inverse_kernel, dim3(1), dim3(256), 0, A, (int)n, W, C, dinfo);
    int *pin = (int *)ctx->pinned;
    DT_HIP(hipMemcpyAsync(pin, dinfo, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    if (pin[0] != 0) return dt_fail(ctx, DT_ENUMERIC, "dt_inverse_f64: singular matrix (column %d)", pin[0]);
    return DT_OK;
}

// ---------------------------------------------------------------------------
// Cyclic Jacobi eigensolver, one workgroup, round-robin parallel ordering.
// H (n x n symmetric) is overwritten; V receives the eigenvectors (columns),
// w the eigenvalues, both sorted in descending order.  Work buffer: perm.
__global__ __launch_bounds__(256) void jacobi_kernel(double *__restrict__ H, int n,
                                                     double *__restrict__ V,
                                                     double *__restrict__ w,
                                                     double *__restrict__ rot,
                                                     int *__restrict__ order) {
    __shared__ double off_s, tot_s;
    __shared__ int stop;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int np = (n + 1) & ~1;  // even number of players (n == np - 1: one bye)
    for (int idx = tid; idx < n * n; idx += nt) V[idx] = (idx % n == idx / n) ? 1.0 : 0.0;
    if (tid == 0) stop = 0;
    __syncthreads();
    // rot layout per pair q: [p, qq, c, s]
    for (int sweep = 0; sweep < 40; sweep++) {
        if (tid == 0) {
            double off = 0.0, tot = 0.0;
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    const double h = H[i + n * j];
                    tot += h * h;
                    if (i != j) off += h * h;
                }
            off_s = off;
            tot_s = tot;
            if (!(off > 1e-30 * tot) || tot == 0.0) stop = 1;
        }
        __syncthreads();
        if (stop) break;
        for (int round = 0; round < np - 1; round++) {
            // pairs of the round-robin schedule: player 0 fixed, others rotate
            for (int q = tid; q < np / 2; q += nt) {
                int a, b;
                if (q == 0) {
                    a = 0;
                    b = 1 + (round % (np - 1));
                } else {
                    a = 1 + ((round + q) % (np - 1));
                    b = 1 + ((round + np - 1 - q) % (np - 1));
                }
                int p = a < b ? a : b, qq = a < b ? b : a;
                double c = 1.0, s = 0.0;
                if (qq < n) {
                    const double apq = H[p + n * qq];
                    if (fabs(apq) > 1e-300) {
                        const double app = H[p + n * p], aqq = H[qq + n * qq];
                        const double tau = (aqq - app) / (2.0 * apq);
                        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                    }
                }
                rot[4 * q + 0] = (double)p;
                rot[4 * q + 1] = (double)qq;
                rot[4 * q + 2] = c;
                rot[4 * q + 3] = s;
            }
            __syncthreads();
            // H <- H J (columns p, q), V <- V J
            for (int idx = tid; idx < (np / 2) * n; idx += nt) {
                const int q = idx / n, r = idx % n;
                const int p = (int)rot[4 * q], qq = (int)rot[4 * q + 1];
                if (qq >= n) continue;
                const double c = rot[4 * q + 2], s = rot[4 * q + 3];
                const double hp = H[r + n * p], hq = H[r + n * qq];
                H[r + n * p] = c * hp - s * hq;
                H[r + n * qq] = s * hp + c * hq;
                const double vp = V[r + n * p], vq = V[r + n * qq];
                V[r + n * p] = c * vp - s * vq;
                V[r + n * qq] = s * vp + c * vq;
            }
            __syncthreads();
            // H <- J^T H (rows p, q)
            for (int idx = tid; idx < (np / 2) * n; idx += nt) {
                const int q = idx / n, r = idx % n;
                const int p = (int)rot[4 * q], qq = (int)rot[4 * q + 1];
                if (qq >= n) continue;
                const double c = rot[4 * q + 2], s = rot[4 * q + 3];
                const double hp = H[p + n * r], hq = H[qq + n * r];
                H[p + n * r] = c * hp - s * hq;
                H[qq + n * r] = s * hp + c * hq;
            }
            __syncthreads();
        }
    }
    // sort descending (selection by one thread; n is small)
    if (tid == 0) {
        for (int i = 0; i < n; i++) order[i] = i;
        for (int i = 0; i < n; i++) {
            int best = i;
            for (int j = i + 1; j < n; j++)
                if (H[order[j] + n * order[j]] > H[order[best] + n * order[best]]) best = j;
            const int t = order[i];
            order[i] = order[best];
            order[best] = t;
        }
        for (int i = 0; i < n; i++) w[i] = H[order[i] + n * order[i]];
    }
    __syncthreads();
    // permute V columns into H (as scratch) then copy back
    for (int idx = tid; idx < n * n; idx += nt) {
        const int r = idx % n, j = idx / n;
        H[idx] = V[r + n * order[j]];
    }
    __syncthreads();
    for (int idx = tid; idx < n * n; idx += nt) V[idx] = H[idx];
}

// Cyclic Jacobi with H and V resident in LDS (n <= LDS_NMAX), same
// round-robin ordering as jacobi_kernel.  H is kept as its upper triangle
// (the input symmetrised on load; element (i, j) at min + ld * max).  A
// round's rotations act on disjoint index pairs, so H <- Jᵀ H J splits into
// 2 x 2 blocks (the rows of pair P, the columns of pair Q) and only the blocks
// with P <= Q are distinct: one thread updates each in place, column rotation
// first and row rotation second.  V <- V J of a round is applied during the
// next round's rotation phase, by waves 1..15 while wave 0 computes the
// rotations (the parameters of consecutive rounds alternate between two
// buffers), and after the last round.  The block and V assignments are
// computed once, so the LDS traffic of a round is the upper triangle and V
// once each.  Two barriers per round (rotations and V, then H).
// One matrix per workgroup of JAC_THREADS: workgroup b solves Hin + b n^2
// into Vout + b n^2, w + b n.  NMAX (<= LDS_NMAX) bounds n: up to 88 every
// thread owns one block of H, above two.
constexpr int JAC_THREADS = 1024;
constexpr int JAC_NSMALL = 88;
template <int NMAX>
__global__ __launch_bounds__(JAC_THREADS) void jacobi_lds_kernel(const double *__restrict__ Hin, int n,
                                                               double *__restrict__ Vout, double *__restrict__ w) {
    Hin += (size_t)blockIdx.x * n * n;
    Vout += (size_t)blockIdx.x * n * n;
    w += (size_t)blockIdx.x * n;
    // odd leading dimension: lanes that walk a row across columns (stride ld
    // doubles) fall on different LDS banks
    __shared__ double H[LDS_NMAX * (LDS_NMAX + 1)];
    __shared__ double V[LDS_NMAX * LDS_NMAX];
    __shared__ double2 rcs[2][LDS_NMAX / 2];  // (c, s) of pair P, rounds alternating
    __shared__ int2 rpq[2][LDS_NMAX / 2];     // (p, q), p < q; q == n is the padding index
    __shared__ double red_off[16], red_dg[16];
    __shared__ int order[LDS_NMAX];
    __shared__ int rotated;  // a rotation of this sweep changed something
    constexpr int NT = JAC_THREADS, NW = NT / 64;
    static_assert(NMAX <= LDS_NMAX, "order above the LDS arrays");
    constexpr int MAXPAIR = (NMAX + 1) / 2;
    constexpr int KH = (MAXPAIR * (MAXPAIR + 1) / 2 + NT - 1) / NT;  // blocks per thread
    constexpr int NVT = NT - 64;                                      // threads of the V update
    constexpr int KV = (NMAX * MAXPAIR + NVT - 1) / NVT;             // V entries per thread
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int np = (n + 1) & ~1, npair = np / 2;  // npair <= 48 < 64
    const int ld = n | 1;
    const int nblk = npair * (npair + 1) / 2, nvit = n * npair;
    // block b = Q (Q + 1) / 2 + P, P <= Q; V entry e = row + n P, e = tid - 64 + NVT k
    int bP[KH], bQ[KH], vr[KV], vP[KV];
#pragma unroll
    for (int k = 0; k < KH; k++) {
        const int b = tid + NT * k < nblk ? tid + NT * k : nblk - 1;
        int Q = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
        while (Q > 0 && Q * (Q + 1) / 2 > b) Q--;
        while ((Q + 1) * (Q + 2) / 2 <= b) Q++;
        bQ[k] = Q;
        bP[k] = b - Q * (Q + 1) / 2;
    }
    const int vt = tid - 64;
#pragma unroll
    for (int k = 0; k < KV; k++) {
        const int e0 = vt + NVT * k;
        const int e = e0 >= 0 && e0 < nvit ? e0 : nvit - 1;
        vP[k] = e / n;
        vr[k] = e - vP[k] * n;
    }
    // V <- V J with the rotations in buffer `buf` (waves 1..15; reads issued together)
    auto apply_v = [&](int buf) {
        int2 vpq[KV];
        double2 vcs[KV];
        double v[KV][2];
#pragma unroll
        for (int k = 0; k < KV; k++) {
            vpq[k] = rpq[buf][vP[k]];
            vcs[k] = rcs[buf][vP[k]];
        }
#pragma unroll
        for (int k = 0; k < KV; k++) {
            const int p1 = vpq[k].y < n ? vpq[k].y : vpq[k].x;
            v[k][0] = V[vr[k] + n * vpq[k].x];
            v[k][1] = V[vr[k] + n * p1];
        }
#pragma unroll
        for (int k = 0; k < KV; k++) {
            if (vt + NVT * k >= nvit) break;
            if (vpq[k].y < n) {  // the padding pair has no column to rotate
                const double c = vcs[k].x, sn = vcs[k].y;
                V[vr[k] + n * vpq[k].x] = c * v[k][0] - sn * v[k][1];
                V[vr[k] + n * vpq[k].y] = sn * v[k][0] + c * v[k][1];
            }
        }
    };
    for (int j = wv; j < n; j += NW)
        for (int i = lane; i <= j; i += 64) H[i + ld * j] = 0.5 * (Hin[i + n * j] + Hin[j + n * i]);
    for (int e = tid; e < n * n; e += NT) V[e] = (e % n) == (e / n) ? 1.0 : 0.0;
    __syncthreads();
    int gr = 0;            // rounds run so far, over all sweeps
    bool pending = false;  // V of round gr - 1 not yet applied
    for (int sweep = 0; sweep < 40; sweep++) {
        double off = 0.0, dg = 0.0;
        for (int j = wv; j < n; j += NW)
            for (int i = lane; i <= j; i += 64) {
                const double h = H[i + ld * j];
                if (i == j)
                    dg += h * h;
                else
                    off += 2.0 * h * h;
            }
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o);
            dg += __shfl_xor(dg, o);
        }
        if (lane == 0) {
            red_off[wv] = off;
            red_dg[wv] = dg;
        }
        __syncthreads();
        off = 0.0;
        dg = 0.0;
        for (int i = 0; i < NW; i++) {
            off += red_off[i];
            dg += red_dg[i];
        }
        __syncthreads();  // red_* reused next sweep
        const double tot = dg + off;
        if (!(off > 1e-30 * tot) || tot == 0.0) break;  // uniform
        if (tid == 0) rotated = 0;
        __syncthreads();
        // thread q < npair of wave 0 pairs a = 1 + (round + q) mod (np - 1) with
        // b = 1 + (round - q) mod (np - 1) (thread 0: 0 with 1 + round), both
        // stepped by one per round without a division
        const int m1 = np - 1;
        int ra = tid % m1, rb = tid == 0 ? 0 : m1 - tid % m1;
        for (int round = 0; round < np - 1; round++, gr++) {
            const int buf = gr & 1;
            if (wv == 0) {
                if (tid < npair) {
                    const int q = tid;
                    const int a = q == 0 ? 0 : 1 + ra;
                    const int b = 1 + rb;
                    ra = ra + 1 == m1 ? 0 : ra + 1;
                    rb = rb + 1 == m1 ? 0 : rb + 1;
                    const int p = a < b ? a : b, qq = a < b ? b : a;
                    double c = 1.0, sn = 0.0;  // identity for the pair holding the padding index n
                    if (qq < n) {
                        const double apq = H[p + ld * qq];
                        const double app = H[p + ld * p], aqq = H[qq + ld * qq];
                        // a rotation is skipped when apq is at the rounding level of the
                        // diagonal pair (a few ulps: rotations leave that much behind)
                        if (fabs(apq) > 1e-300 && fabs(apq) > 1e-15 * sqrt(fabs(app * aqq))) {
                            rotated = 1;
                            const double tau = (aqq - app) / (2.0 * apq);
                            const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                            c = 1.0 / sqrt(1.0 + t * t);
                            sn = t * c;
                        }
                    }
                    rpq[buf][q] = make_int2(p, qq);
                    rcs[buf][q] = make_double2(c, sn);
                }
            } else if (pending) {
                apply_v(buf ^ 1);
            }
            __syncthreads();
            // H: every LDS read of the blocks is issued before the arithmetic;
            // blocks past the thread's share read valid addresses and store nothing
            int2 pa[KH], pb[KH];
            double2 ca[KH], cb[KH];
#pragma unroll
            for (int k = 0; k < KH; k++) {
                pa[k] = rpq[buf][bP[k]];
                pb[k] = rpq[buf][bQ[k]];
                ca[k] = rcs[buf][bP[k]];
                cb[k] = rcs[buf][bQ[k]];
            }
            int adr[KH][4];
            double h[KH][4];
#pragma unroll
            for (int k = 0; k < KH; k++) {
                const int p0 = pa[k].x, p1 = pa[k].y, q0 = pb[k].x, q1 = pb[k].y;
                const bool r1 = p1 < n, k1 = q1 < n;
                auto at = [ld](int i, int j) { return i <= j ? i + ld * j : j + ld * i; };
                adr[k][0] = at(p0, q0);
                adr[k][1] = k1 ? at(p0, q1) : adr[k][0];
                adr[k][2] = r1 ? at(p1, q0) : adr[k][0];
                adr[k][3] = (r1 && k1) ? at(p1, q1) : adr[k][0];
#pragma unroll
                for (int u = 0; u < 4; u++) h[k][u] = H[adr[k][u]];
            }
#pragma unroll
            for (int k = 0; k < KH; k++) {
                if (tid + NT * k >= nblk) break;
                const bool r1 = pa[k].y < n, k1 = pb[k].y < n, diag = bP[k] == bQ[k];
                const double cP = ca[k].x, sP = ca[k].y, cQ = cb[k].x, sQ = cb[k].y;
                const double h00 = h[k][0], h01 = k1 ? h[k][1] : 0.0;
                const double h10 = r1 ? h[k][2] : 0.0, h11 = (r1 && k1) ? h[k][3] : 0.0;
                // columns: (x0, x1) = (cQ h0 - sQ h1, sQ h0 + cQ h1) per row
                const double x00 = cQ * h00 - sQ * h01, x01 = sQ * h00 + cQ * h01;
                const double x10 = cQ * h10 - sQ * h11, x11 = sQ * h10 + cQ * h11;
                // rows: (y0, y1) = (cP x0 - sP x1, sP x0 + cP x1) per column
                H[adr[k][0]] = cP * x00 - sP * x10;
                if (k1) H[adr[k][1]] = cP * x01 - sP * x11;
                // a diagonal block's (p1, p0) is its (p0, p1), stored once
                if (r1 && !diag) H[adr[k][2]] = sP * x00 + cP * x10;
                if (r1 && k1) H[adr[k][3]] = sP * x01 + cP * x11;
            }
            pending = true;
            __syncthreads();
        }
        if (!rotated) break;  // uniform: read after the round's barrier
    }
    if (pending) {
        if (wv > 0) apply_v((gr - 1) & 1);
        __syncthreads();
    }
    // descending order by rank (ties broken by index)
    for (int i = tid; i < n; i += NT) {
        const double di = H[i + ld * i];
        int rank = 0;
        for (int j = 0; j < n; j++) {
            const double dj = H[j + ld * j];
            rank += (dj > di) || (dj == di && j < i);
        }
        order[rank] = i;
    }
    __syncthreads();
    for (int i = tid; i < n; i += NT) w[i] = H[order[i] + ld * order[i]];
    for (int idx = tid; idx < n * n; idx += NT) {
        const int r = idx % n, j = idx / n;
        Vout[idx] = V[r + n * order[j]];
    }
}

// `count` independent n x n problems (n <= LDS_NMAX) stored back to back, one
// workgroup each.
static int launch_jacobi_lds(dt_ctx *ctx, double *H, size_t n, int count, double *V, double *w) {
    if (n <= (size_t)JAC_NSMALL) {
        DT_LAUNCH("jacobi", jacobi_lds_kernel<JAC_NSMALL>, dim3((unsigned)count), dim3(JAC_THREADS), 0, H, (int)n,
                  V, w);
    } else {
        DT_LAUNCH("jacobi", jacobi_lds_kernel<LDS_NMAX>, dim3((unsigned)count), dim3(JAC_THREADS), 0, H, (int)n, V,
                  w);
    }
    return DT_OK;
}

int dt_jacobi_eig_batched(dt_ctx *ctx, double *H, size_t n, int count, double *V, double *w) {
    if (n > (size_t)LDS_NMAX || count < 1) return dt_fail(ctx, DT_EUNSUPPORTED, "batched Jacobi: n %zu", n);
    return launch_jacobi_lds(ctx, H, n, count, V, w);
}

int dt_jacobi_eig(dt_ctx *ctx, double *H, size_t n, double *V, double *w) {
    const size_t np = (n + 1) & ~(size_t)1;
    double *rot = (double *)dt_ws(ctx, WS_EIG_H, (2 * np + 8) * sizeof(double) + n * sizeof(int) + 256);
    if (!rot) return DT_ENOMEM;
    int *order = (int *)(rot + 2 * np + 8);
    if (n <= (size_t)LDS_NMAX) {
        return launch_jacobi_lds(ctx, H, n, 1, V, w);
    }
    DT_LAUNCH("jacobi", jacobi_kernel, dim3(1), dim3(256), 0, H, (int)n, V, w, rot, order);
    return DT_OK;
}

// ---------------------------------------------------------------------------
// Leading r eigenpairs of a symmetric n x n FP64 matrix.
extern "C" int dt_syev_top_f64(dt_ctx *ctx, const double *S, size_t n, size_t r, double *V,
                               double *w) {
    DT_CHECK_ARG(ctx && S && V && r >= 1 && r <= n, "dt_syev_top_f64: bad argument");
    if (n <= (size_t)LDS_NMAX) {
        double *H = (double *)dt_ws(ctx, WS_EIG_X, 2 * n * n * sizeof(double) + n * sizeof(double));
        if (!H) return DT_ENOMEM;
        double *Vall = H + n * n;
        double *wall = Vall + n * n;
        DT_HIP(hipMemcpyAsync(H, S, n * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        DT_TRY(dt_jacobi_eig(ctx, H, n, Vall, wall));
        DT_HIP(hipMemcpyAsync(V, Vall, n * r * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        if (w) DT_HIP(hipMemcpyAsync(w, wall, r * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        return DT_OK;
    }
    // Oversampled subspace iteration with Rayleigh-Ritz every few steps.
    // oversampling: r + 16 keeps the Rayleigh-Ritz problem in LDS for r <= 80
    const size_t p = std::min(n, r + 16 <= (size_t)LDS_NMAX ? r + 16 : std::max(2 * r, r + 16));
    double *X = (double *)dt_ws(ctx, WS_EIG_X, (2 * n * p + 2 * p * p + 2 * p) * sizeof(double));
    if (!X) return DT_ENOMEM;
    double *Y = X + n * p;
    double *H = Y + n * p;
    double *W = H + p * p;
    double *theta = W + p * p;
    DT_TRY(dt_fill_normal_f64(ctx, X, n * p, 0x5eed5eedULL, 77));
    DT_TRY(dt_qr_f64(ctx, X, nullptr, X, n, p));
    std::vector<double> prev(r, 0.0), cur(p);
    const int max_iter = 600;
    for (int it = 1; it <= max_iter; it++) {
        DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, n, p, n, 1.0, S, n, X, n, 0.0, Y, n));
        if (it % 4 == 0 || it == max_iter) {
            // Rayleigh-Ritz on span(X): H = X^T S X
            DT_TRY(dt_gemm_f64_internal(ctx, 1, 0, p, p, n, 1.0, X, n, Y, n, 0.0, H, p));
            DT_TRY(dt_jacobi_eig(ctx, H, p, W, theta));
            DT_HIP(hipMemcpyAsync(cur.data(), theta, p * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            DT_HIP(hipStreamSynchronize(ctx->stream));
            double diff = 0.0;
            for (size_t i = 0; i < r; i++) diff = std::max(diff, std::fabs(cur[i] - prev[i]));
            for (size_t i = 0; i < r; i++) prev[i] = cur[i];
            const double scale = std::max(std::fabs(cur[0]), 1e-300);
            if (diff <= 1e-14 * scale || it == max_iter) {
                // V = X W[:, :r]  (Ritz vectors)
                DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, n, r, p, 1.0, X, n, W, p, 0.0, V, n));
                if (w) DT_HIP(hipMemcpyAsync(w, theta, r * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
                return DT_OK;
            }
        }
        DT_TRY(dt_qr_f64(ctx, X, nullptr, Y, n, p));
    }
    return DT_OK;
}
