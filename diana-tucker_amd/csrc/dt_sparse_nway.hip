// Sparse TTMc for any number of modes and ranks, FP32 or FP64 values.
//
// Reference: Function::ttmc_mTR / ttmc_mTL (template/function/sp_tensor.tpp:
// 236-365) accumulate, per mode-n slice, the Kronecker rows of the other
// modes' factor rows (add_outer_product_all, :367-417: the first other mode
// fastest), then contract them with Ut_n (mTR: L = Y_(n)ᵀ U_n) or with L
// (mTL: Ut_n = (Y_(n) L)ᵀ).  The 3-way rank-16 FP32 case has its own MFMA
// pipeline (dt_sparse.hip); this file serves every other shape and the FP64
// values of SpTensor<double>.
//
// spg_ttmc: the mode-sorted nonzero stream is cut into equal chunks, one
// workgroup each.  K = R_a * M, with a the first other mode and M the product
// of the rest; thread t owns the entries k = t + 256 q.  A batch of B nonzeros
// is staged in LDS: their factor rows (every other mode) and, per nonzero,
// the rest-product P[j] = val * prod_{m != n, a} U_m[i_m, r_m(j)] (M entries,
// formed cooperatively).  Then, nonzero by nonzero in stream order, every
// thread adds U_a[i_a, r_a(k)] * P[j(k)] to its entries; a row change flushes
// the row -- into Y_c when the chunk holds all of the row, else into one of
// the chunk's two carry slots (row begun in an earlier chunk: slot 0; row
// continuing into the next chunk: slot 1).  spg_fixup adds, for each row
// begun in chunk c and continuing, slot 1 of c and slot 0 of the following
// chunks that hold the row, in chunk order (deterministic).
#include "dt_sparse.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace {

constexpr int G_THREADS = 256;
constexpr int G_NQ = DT_SPG_KMAX / G_THREADS;  // entries per thread
constexpr int G_LDS = 48 * 1024;               // staging budget of spg_ttmc

template <typename T>
struct SpgArgs {
    const int32_t *perm, *row, *slot;
    const int32_t *idx[DT_MAX_NDIM - 1];  // the other modes, ascending (input order)
    const T *U[DT_MAX_NDIM - 1];          // their factor rows (I_m x R_m row-major)
    int R[DT_MAX_NDIM - 1];
    int off[DT_MAX_NDIM - 1];  // offset of mode m's row in a staged block
    int nother;
    const T *val;
    int64_t nnz;
    int K, M, S;  // K = R_a M; S = sum of the other ranks
    int B;        // nonzeros per staged batch
    int64_t chunk, nchunks;
};

template <typename T>
__global__ __launch_bounds__(G_THREADS) void spg_ttmc_kernel(SpgArgs<T> a, T *__restrict__ Yc, T *__restrict__ carry,
                                                             int32_t *__restrict__ crow) {
    extern __shared__ double g_lds[];
    T *F = (T *)g_lds;                   // B x S factor rows
    T *P = F + (size_t)a.B * a.S;        // B x M rest-products
    int32_t *RW = (int32_t *)(P + (size_t)a.B * a.M);  // B sorted rows
    const int tid = threadIdx.x;
    const int64_t c = blockIdx.x;
    const int64_t p0 = c * a.chunk;
    const int64_t p1 = std::min<int64_t>(p0 + a.chunk, a.nnz);
    const int Ra = a.R[0];
    const int row_first = a.row[p0], row_last = a.row[p1 - 1];
    const bool first_cut = p0 > 0 && a.row[p0 - 1] == row_first;
    const bool last_cut = p1 < a.nnz && a.row[p1] == row_last;
    if (tid == 0) {
        crow[2 * c] = first_cut ? row_first : -1;
        crow[2 * c + 1] = (last_cut && !(first_cut && row_first == row_last)) ? row_last : -1;
    }
    // entry q of this thread: k = tid + 256 q = ra + Ra j
    int ra[G_NQ], jj[G_NQ];
#pragma unroll
    for (int q = 0; q < G_NQ; q++) {
        const int k = tid + G_THREADS * q;
        ra[q] = k < a.K ? k % Ra : 0;
        jj[q] = k < a.K ? k / Ra : 0;
    }
    T acc[G_NQ];
#pragma unroll
    for (int q = 0; q < G_NQ; q++) acc[q] = (T)0;
    int cur = row_first;
    auto flush = [&](int r, bool at_end) {
        T *dst;
        if (r == row_first && first_cut) {
            dst = carry + (size_t)(2 * c) * a.K;
        } else if (at_end && last_cut) {
            dst = carry + (size_t)(2 * c + 1) * a.K;
        } else {
            dst = Yc + (size_t)a.slot[r] * a.K;
        }
#pragma unroll
        for (int q = 0; q < G_NQ; q++) {
            const int k = tid + G_THREADS * q;
            if (k < a.K) dst[k] = acc[q];
            acc[q] = (T)0;
        }
    };
    for (int64_t b0 = p0; b0 < p1; b0 += a.B) {
        const int nb = (int)std::min<int64_t>(a.B, p1 - b0);
        // stage the factor rows and the sorted rows of the batch
        for (int e = tid; e < nb * a.S; e += G_THREADS) {
            const int b = e / a.S, o = e % a.S;
            int m = 0;
            while (m + 1 < a.nother && o >= a.off[m + 1]) m++;
            const int32_t src = a.perm[b0 + b];
            F[e] = a.U[m][(size_t)a.idx[m][src] * a.R[m] + (o - a.off[m])];
        }
        for (int b = tid; b < nb; b += G_THREADS) RW[b] = a.row[b0 + b];
        __syncthreads();
        // rest-products: P[b][j] = val * prod over the modes after a
        for (int e = tid; e < nb * a.M; e += G_THREADS) {
            const int b = e / a.M;
            int j = e % a.M;
            T v = a.val[a.perm[b0 + b]];
            for (int m = 1; m < a.nother; m++) {
                const int r = j % a.R[m];
                j /= a.R[m];
                v *= F[(size_t)b * a.S + a.off[m] + r];
            }
            P[e] = v;
        }
        __syncthreads();
        for (int b = 0; b < nb; b++) {
            const int r = RW[b];
            if (r != cur) {
                flush(cur, false);
                cur = r;
            }
            const T *Fb = F + (size_t)b * a.S;  // mode a's row first (off[0] = 0)
            const T *Pb = P + (size_t)b * a.M;
#pragma unroll
            for (int q = 0; q < G_NQ; q++)
                if (tid + G_THREADS * q < a.K) acc[q] += Fb[ra[q]] * Pb[jj[q]];
        }
        __syncthreads();
    }
    flush(cur, true);
}

// row begun in chunk c and continuing: slot 1 of c plus slot 0 of the next
// chunks that hold it, in chunk order
template <typename T>
__global__ __launch_bounds__(G_THREADS) void spg_fixup(const T *__restrict__ carry, const int32_t *__restrict__ crow,
                                                       int64_t nchunks, int K, const int32_t *__restrict__ slot,
                                                       T *__restrict__ Yc) {
    const int64_t c = blockIdx.x;
    const int r = crow[2 * c + 1];
    if (r < 0) return;
    for (int k = threadIdx.x; k < K; k += G_THREADS) {
        T s = carry[(size_t)(2 * c + 1) * K + k];
        for (int64_t c2 = c + 1; c2 < nchunks && crow[2 * c2] == r; c2++) s += carry[(size_t)(2 * c2) * K + k];
        Yc[(size_t)slot[r] * K + k] = s;
    }
}

// part[split][k + K j] = sum over the split's compact rows s of Y_c[s][k] U[rows[s]][j]
// workgroup = 64 entries of k x all R columns (thread: k = k0 + tid % 64,
// j = tid / 64 + 4 q)
template <typename T>
__global__ __launch_bounds__(G_THREADS) void spg_ytu_kernel(const T *__restrict__ Yc, const int32_t *__restrict__ rows,
                                                            int64_t nrows, const T *__restrict__ Un, int K, int R,
                                                            int64_t rows_per_split, double *__restrict__ part) {
    constexpr int SB = 16;
    __shared__ T Ys[SB][64];
    __shared__ T Us[SB][DT_SPG_RMAX];
    const int tid = threadIdx.x, kk = tid & 63, jg = tid >> 6;
    const int k0 = blockIdx.x * 64;
    const int64_t s0 = (int64_t)blockIdx.y * rows_per_split;
    const int64_t s1 = std::min<int64_t>(s0 + rows_per_split, nrows);
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = 0.0;
    for (int64_t sb = s0; sb < s1; sb += SB) {
        const int ns = (int)std::min<int64_t>(SB, s1 - sb);
        for (int e = tid; e < SB * 64; e += G_THREADS) {
            const int s = e / 64, k = e % 64;
            Ys[s][k] = (s < ns && k0 + k < K) ? Yc[(size_t)(sb + s) * K + k0 + k] : (T)0;
        }
        for (int e = tid; e < SB * DT_SPG_RMAX; e += G_THREADS) {
            const int s = e / DT_SPG_RMAX, j = e % DT_SPG_RMAX;
            Us[s][j] = (s < ns && j < R) ? Un[(size_t)rows[sb + s] * R + j] : (T)0;
        }
        __syncthreads();
        for (int s = 0; s < ns; s++) {
            const double y = (double)Ys[s][kk];
#pragma unroll
            for (int q = 0; q < 16; q++) acc[q] += y * (double)Us[s][jg + 4 * q];
        }
        __syncthreads();
    }
    const int k = k0 + kk;
    if (k < K) {
        double *dst = part + (size_t)blockIdx.y * K * R;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int j = jg + 4 * q;
            if (j < R) dst[k + (size_t)K * j] = acc[q];
        }
    }
}

__global__ void spg_reduce(const double *__restrict__ part, int64_t ne, int nsplit, double *__restrict__ out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int p = 0; p < nsplit; p++) s += part[(size_t)p * ne + e];
        out[e] = s;
    }
}

// W[rows[s] * sr + j * sc] = sum_k Y_c[s][k] L[k + K j], 64 compact rows per workgroup
template <typename T>
__global__ __launch_bounds__(G_THREADS) void spg_yl_kernel(const T *__restrict__ Yc, const int32_t *__restrict__ rows,
                                                           int64_t nrows, const double *__restrict__ L, int K, int R,
                                                           double *__restrict__ W, int64_t sr, int64_t sc) {
    constexpr int KC = 32;  // k per staged chunk
    __shared__ double Ys[64][KC + 1];
    __shared__ double Ls[KC][DT_SPG_RMAX + 1];
    const int tid = threadIdx.x, rr = tid & 63, jg = tid >> 6;
    const int64_t s0 = (int64_t)blockIdx.x * 64;
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = 0.0;
    for (int k0 = 0; k0 < K; k0 += KC) {
        for (int e = tid; e < 64 * KC; e += G_THREADS) {
            const int s = e / KC, k = e % KC;
            Ys[s][k] = (s0 + s < nrows && k0 + k < K) ? (double)Yc[(size_t)(s0 + s) * K + k0 + k] : 0.0;
        }
        for (int e = tid; e < KC * DT_SPG_RMAX; e += G_THREADS) {
            const int k = e / DT_SPG_RMAX, j = e % DT_SPG_RMAX;
            Ls[k][j] = (k0 + k < K && j < R) ? L[(k0 + k) + (size_t)K * j] : 0.0;
        }
        __syncthreads();
        for (int k = 0; k < KC; k++) {
            const double y = Ys[rr][k];
#pragma unroll
            for (int q = 0; q < 16; q++) acc[q] += y * Ls[k][jg + 4 * q];
        }
        __syncthreads();
    }
    if (s0 + rr < nrows) {
        const int64_t i = rows[s0 + rr];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int j = jg + 4 * q;
            if (j < R) W[i * sr + (int64_t)j * sc] = acc[q];
        }
    }
}

template <typename T>
__global__ void spg_rows_view_kernel(const double *__restrict__ U, int64_t I, int R, T *__restrict__ Ur) {
    const int64_t n = I * R;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / R, r = e % R;
        Ur[e] = (T)U[i + I * r];
    }
}

__global__ __launch_bounds__(256) void spg_norm_kernel(const double *__restrict__ v, int64_t n, double *__restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += v[i] * v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sqrt(red[0]);
}

inline dim3 g_grid(int64_t n) { return dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65535)); }

template <typename T>
int ttmc_t(dt_ctx *ctx, const dt_sp *sp, int n, const void *const *Urow, const int *R, T *Yc) {
    const dt_sp_mode &md = sp->mode[n];
    SpgArgs<T> a{};
    a.perm = md.perm;
    a.row = md.row;
    a.slot = md.slot;
    a.nother = 0;
    int64_t K = 1;
    int S = 0;
    for (int m = 0; m < sp->ndim; m++) {
        if (m == n) continue;
        DT_CHECK_ARG(R[m] >= 1 && R[m] <= DT_SPG_RMAX, "sparse TTMc: rank %d of mode %d not in [1, %d]", R[m], m,
                     DT_SPG_RMAX);
        a.idx[a.nother] = sp->idx[m];
        a.U[a.nother] = (const T *)Urow[m];
        a.R[a.nother] = R[m];
        a.off[a.nother] = S;
        S += R[m];
        K *= R[m];
        a.nother++;
    }
    DT_CHECK_ARG(K <= DT_SPG_KMAX, "sparse TTMc: prod of the other ranks %lld > %d", (long long)K, DT_SPG_KMAX);
    a.val = std::is_same<T, double>::value ? (const T *)(const void *)sp->val64 : (const T *)(const void *)sp->val;
    DT_CHECK_ARG(a.val, "sparse TTMc: the tensor has no FP64 values");
    a.nnz = sp->nnz;
    a.K = (int)K;
    a.M = (int)(K / a.R[0]);
    a.S = S;
    const size_t per_nz = (size_t)(a.S + a.M) * sizeof(T) + 4;
    a.B = (int)std::max<size_t>(1, std::min<size_t>(64, G_LDS / per_nz));
    // chunks: enough workgroups to fill the GPU, carries within 2^26 elements
    const int64_t by_mem = std::max<int64_t>(1, ((int64_t)1 << 26) / (2 * K));
    const int64_t want = std::min<int64_t>(16384, by_mem);
    a.chunk = std::max<int64_t>(64, ceil_div(sp->nnz, want));
    a.nchunks = ceil_div(sp->nnz, a.chunk);
    char *ws = (char *)dt_ws(ctx, WS_SPARSE0, (size_t)a.nchunks * (2 * K * sizeof(T) + 8) + 512);
    if (!ws) return DT_ENOMEM;
    T *carry = (T *)ws;
    int32_t *crow = (int32_t *)(ws + (size_t)a.nchunks * 2 * K * sizeof(T));
    const size_t lds = (size_t)a.B * per_nz + 64;
    DT_LAUNCH("spg_ttmc", spg_ttmc_kernel<T>, dim3((unsigned)a.nchunks), dim3(G_THREADS), lds, a, Yc, carry, crow);
    DT_LAUNCH("spg_fixup", spg_fixup<T>, dim3((unsigned)a.nchunks), dim3(G_THREADS), 0, (const T *)carry,
              (const int32_t *)crow, a.nchunks, a.K, md.slot, Yc);
    return DT_OK;
}

template <typename T>
int ytu_t(dt_ctx *ctx, const dt_sp *sp, int n, const T *Un, const int *R, const T *Yc, double *L) {
    const dt_sp_mode &md = sp->mode[n];
    int64_t K = 1;
    for (int m = 0; m < sp->ndim; m++)
        if (m != n) K *= R[m];
    const int Rn = R[n];
    DT_CHECK_ARG(Rn >= 1 && Rn <= DT_SPG_RMAX, "sparse L: rank %d not in [1, %d]", Rn, DT_SPG_RMAX);
    const int64_t nk = ceil_div(K, 64);
    const int64_t by_mem = std::max<int64_t>(1, ((int64_t)1 << 24) / (K * Rn));
    int64_t nsplit = std::min<int64_t>(std::max<int64_t>(1, 1024 / nk), by_mem);
    nsplit = std::max<int64_t>(1, std::min<int64_t>(nsplit, ceil_div(std::max<int64_t>(md.nrows, 1), 64)));
    const int64_t rps = ceil_div(std::max<int64_t>(md.nrows, 1), nsplit);
    nsplit = ceil_div(std::max<int64_t>(md.nrows, 1), rps);
    double *part = (double *)dt_ws(ctx, WS_SPARSE1, (size_t)nsplit * K * Rn * 8 + 256);
    if (!part) return DT_ENOMEM;
    DT_LAUNCH("spg_ytu", spg_ytu_kernel<T>, dim3((unsigned)nk, (unsigned)nsplit), dim3(G_THREADS), 0, Yc, md.rows,
              md.nrows, Un, (int)K, Rn, rps, part);
    DT_LAUNCH("spg_ytu", spg_reduce, g_grid(K * Rn), dim3(256), 0, (const double *)part, K * Rn, (int)nsplit, L);
    return DT_OK;
}

template <typename T>
int yl_t(dt_ctx *ctx, const dt_sp *sp, int n, const int *R, const T *Yc, const double *L, double *W, bool rowmajor) {
    const dt_sp_mode &md = sp->mode[n];
    int64_t K = 1;
    for (int m = 0; m < sp->ndim; m++)
        if (m != n) K *= R[m];
    const int Rn = R[n];
    DT_CHECK_ARG(Rn >= 1 && Rn <= DT_SPG_RMAX, "sparse Y L: rank %d not in [1, %d]", Rn, DT_SPG_RMAX);
    const int64_t I = sp->dims[n];
    DT_TRY(dt_zero(ctx, W, (size_t)I * Rn * 8));
    if (md.nrows == 0) return DT_OK;
    const int64_t sr = rowmajor ? Rn : 1, sc = rowmajor ? 1 : I;
    DT_LAUNCH("spg_yl", spg_yl_kernel<T>, dim3((unsigned)ceil_div(md.nrows, 64)), dim3(G_THREADS), 0, Yc, md.rows,
              md.nrows, L, (int)K, Rn, W, sr, sc);
    return DT_OK;
}

}  // namespace

int dt_spg_ttmc(dt_ctx *ctx, const dt_sp *sp, int n, const void *const *Urow, int T, const int *R, void *Yc) {
    if (T == DT_F64) return ttmc_t<double>(ctx, sp, n, Urow, R, (double *)Yc);
    return ttmc_t<float>(ctx, sp, n, Urow, R, (float *)Yc);
}

int dt_spg_ytu(dt_ctx *ctx, const dt_sp *sp, int n, const void *Un_row, int T, const int *R, const void *Yc,
               double *L) {
    if (T == DT_F64) return ytu_t<double>(ctx, sp, n, (const double *)Un_row, R, (const double *)Yc, L);
    return ytu_t<float>(ctx, sp, n, (const float *)Un_row, R, (const float *)Yc, L);
}

int dt_spg_yl(dt_ctx *ctx, const dt_sp *sp, int n, int T, const int *R, const void *Yc, const double *L, double *W) {
    if (T == DT_F64) return yl_t<double>(ctx, sp, n, R, (const double *)Yc, L, W, false);
    return yl_t<float>(ctx, sp, n, R, (const float *)Yc, L, W, false);
}

int dt_spg_rows_view(dt_ctx *ctx, const double *U64, int64_t I, int R, void *Urow, int T) {
    if (T == DT_F64) {
        DT_LAUNCH("spg_view", spg_rows_view_kernel<double>, g_grid(I * R), dim3(256), 0, U64, I, R, (double *)Urow);
    } else {
        DT_LAUNCH("spg_view", spg_rows_view_kernel<float>, g_grid(I * R), dim3(256), 0, U64, I, R, (float *)Urow);
    }
    return DT_OK;
}

int dt_spg_norm(dt_ctx *ctx, const double *v, int64_t n, double *out) {
    DT_LAUNCH("spg_norm", spg_norm_kernel, dim3(1), dim3(256), 0, v, n, out);
    return DT_OK;
}

// ---- Function:: operators at the C ABI ------------------------------------------
namespace {
// element type of a tensor's generic path, its row views of Ut (R_m x I_m
// column-major FP64 = I_m x R_m row-major) and Y_c
struct SpgOperands {
    int T = DT_F32;
    const void *Urow[DT_MAX_NDIM] = {};
    int R[DT_MAX_NDIM] = {};
    void *Yc = nullptr;
    int64_t K = 1;
};

int spg_operands(dt_ctx *ctx, const dt_sp *sp, int n, const double *const *Ut, const size_t *ranks, SpgOperands &o,
                 const char *who) {
    DT_CHECK_ARG(ctx && sp && Ut && ranks, "%s: null argument", who);
    DT_CHECK_ARG(n >= 0 && n < sp->ndim, "%s: bad mode %d", who, n);
    o.T = sp->val64 ? DT_F64 : DT_F32;
    const size_t esz = o.T == DT_F64 ? 8 : 4;
    size_t view_elems = 0;
    for (int m = 0; m < sp->ndim; m++) {
        DT_CHECK_ARG(ranks[m] >= 1 && ranks[m] <= (size_t)DT_SPG_RMAX && Ut[m],
                     "%s: rank %zu of mode %d not in [1, %d] (or factor missing)", who, ranks[m], m, DT_SPG_RMAX);
        o.R[m] = (int)ranks[m];
        if (m != n) o.K *= o.R[m];
        if (o.T == DT_F32) view_elems += (size_t)sp->dims[m] * ranks[m];
    }
    DT_CHECK_ARG(o.K <= DT_SPG_KMAX, "%s: prod of the other ranks %lld > %d", who, (long long)o.K, DT_SPG_KMAX);
    const size_t yc_bytes = (size_t)std::max<int64_t>(sp->mode[n].nrows, 1) * o.K * esz;
    char *ws = (char *)dt_ws(ctx, WS_SPARSE2, yc_bytes + view_elems * 4 + 256 * (DT_MAX_NDIM + 1));
    if (!ws) return DT_ENOMEM;
    o.Yc = ws;
    char *v = ws + ((yc_bytes + 255) & ~(size_t)255);
    for (int m = 0; m < sp->ndim; m++) {
        if (o.T == DT_F64) {
            o.Urow[m] = Ut[m];  // already I_m x R_m row-major
        } else {
            // FP32 rows of Ut (same layout)
            DT_TRY(dt_convert(ctx, v, DT_F32, Ut[m], DT_F64, (size_t)sp->dims[m] * ranks[m]));
            o.Urow[m] = v;
            v += ((size_t)sp->dims[m] * ranks[m] * 4 + 255) & ~(size_t)255;
        }
    }
    return DT_OK;
}
}  // namespace

extern "C" int dt_sp_ttmc_mtr(dt_ctx *ctx, dt_sp *sp, int n, const double *const *Ut, const size_t *ranks,
                              double *L) {
    SpgOperands o;
    DT_TRY(spg_operands(ctx, sp, n, Ut, ranks, o, "dt_sp_ttmc_mtr"));
    DT_CHECK_ARG(L, "dt_sp_ttmc_mtr: null L");
    DT_TRY(dt_spg_ttmc(ctx, sp, n, o.Urow, o.T, o.R, o.Yc));
    return dt_spg_ytu(ctx, sp, n, o.Urow[n], o.T, o.R, o.Yc, L);
}

extern "C" int dt_sp_ttmc_mtl(dt_ctx *ctx, dt_sp *sp, int n, const double *const *Ut, const size_t *ranks,
                              const double *L, double *ret) {
    SpgOperands o;
    DT_TRY(spg_operands(ctx, sp, n, Ut, ranks, o, "dt_sp_ttmc_mtl"));
    DT_CHECK_ARG(L && ret, "dt_sp_ttmc_mtl: null L or output");
    DT_TRY(dt_spg_ttmc(ctx, sp, n, o.Urow, o.T, o.R, o.Yc));
    // ret (R_n x I_n column-major) = (Y_(n) L)ᵀ, i.e. Y_(n) L stored row-major
    if (o.T == DT_F64) return yl_t<double>(ctx, sp, n, o.R, (const double *)o.Yc, L, ret, true);
    return yl_t<float>(ctx, sp, n, o.R, (const float *)o.Yc, L, ret, true);
}

// ---- duplicate coordinates merged at ingest ----------------------------------------
// The COO is sorted stably by every mode in turn (lowest mode first, so the
// last pass's mode is the most significant): equal coordinates become
// adjacent, in input order.  Each run's values are summed in FP64 in input
// order and kept at the run's first input position; the survivors are then
// compacted in input order, so a file without duplicates comes out unchanged.
namespace {
__global__ void mg_iota(int32_t *__restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int32_t)i;
}
__global__ void mg_keys(const int32_t *__restrict__ idx, const int32_t *__restrict__ perm, int64_t n,
                        uint32_t *__restrict__ key) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        key[i] = (uint32_t)idx[perm[i]];
}
__global__ void mg_check(const int32_t *__restrict__ idx, int64_t n, int64_t dim, int32_t *__restrict__ flag) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (idx[i] < 0 || idx[i] >= dim) atomicOr(flag, 1);
}
struct MgIdx {
    int32_t *idx[DT_MAX_NDIM];
    int32_t *out[DT_MAX_NDIM];
    int ndim;
};
// head[p] = 1 where the sorted coordinate differs from its predecessor
__global__ void mg_heads(MgIdx a, const int32_t *__restrict__ perm, int64_t n, int32_t *__restrict__ head) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        int h = p == 0;
        if (!h) {
            const int32_t x = perm[p], y = perm[p - 1];
            for (int m = 0; m < a.ndim; m++) h |= a.idx[m][x] != a.idx[m][y];
        }
        head[p] = h;
    }
}
// a run's FP64 sum in input order, kept at its first input position
__global__ void mg_sums(const int32_t *__restrict__ perm, const int32_t *__restrict__ head, int64_t n,
                        const double *__restrict__ val, int32_t *__restrict__ keep, double *__restrict__ vsum) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        if (!head[p]) continue;
        double s = val[perm[p]];
        for (int64_t q = p + 1; q < n && !head[q]; q++) s += val[perm[q]];
        keep[perm[p]] = 1;
        vsum[perm[p]] = s;
    }
}
__global__ void mg_compact(MgIdx a, const int32_t *__restrict__ keep, const int32_t *__restrict__ pos, int64_t n,
                           const double *__restrict__ vsum, double *__restrict__ vout) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!keep[i]) continue;
        const int32_t j = pos[i];
        for (int m = 0; m < a.ndim; m++) a.out[m][j] = a.idx[m][i];
        vout[j] = vsum[i];
    }
}
int mg_bits(int64_t dim) {
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) < dim) b++;
    return b;
}
}  // namespace

extern "C" int dt_sp_merge_duplicates(dt_ctx *ctx, int ndim, const size_t *dims, int64_t nnz, int32_t *const *idx,
                                      double *val, int64_t *nnz_out) {
    DT_CHECK_ARG(ctx && dims && idx && val && nnz_out, "dt_sp_merge_duplicates: null argument");
    DT_CHECK_ARG(ndim >= 1 && ndim <= DT_MAX_NDIM, "dt_sp_merge_duplicates: bad ndim %d", ndim);
    DT_CHECK_ARG(nnz >= 1 && nnz < ((int64_t)1 << 31), "dt_sp_merge_duplicates: nnz out of range");
    for (int m = 0; m < ndim; m++)
        DT_CHECK_ARG(idx[m] && dims[m] >= 1 && dims[m] < ((size_t)1 << 31), "dt_sp_merge_duplicates: mode %d", m);
    const size_t n = (size_t)nnz;
    const size_t al = 256;
    auto up = [&](size_t b) { return (b + al - 1) & ~(al - 1); };
    const size_t sort_tmp = std::max(dt_radix_sort_tmp_bytes(nnz, 4), dt_scan_tmp_bytes(nnz));
    const size_t bytes = 2 * up(n * 4) /*perm*/ + 2 * up(n * 4) /*keys*/ + up(sort_tmp) + 3 * up(n * 4) /*head keep pos*/ +
                         2 * up(n * 8) /*vsum vout*/ + (size_t)ndim * up(n * 4) /*idx out*/ + 512;
    void *buf = nullptr;
    if (hipMalloc(&buf, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return dt_fail(ctx, DT_ENOMEM, "dt_sp_merge_duplicates: hipMalloc(%zu) failed", bytes);
    }
    char *c = (char *)buf;
    auto take = [&](size_t b) {
        char *r = c;
        c += up(b);
        return r;
    };
    int32_t *perm = (int32_t *)take(n * 4), *perm2 = (int32_t *)take(n * 4);
    uint32_t *key = (uint32_t *)take(n * 4), *key2 = (uint32_t *)take(n * 4);
    void *stmp = take(sort_tmp);
    int32_t *head = (int32_t *)take(n * 4), *keep = (int32_t *)take(n * 4), *pos = (int32_t *)take(n * 4);
    double *vsum = (double *)take(n * 8), *vout = (double *)take(n * 8);
    MgIdx a{};
    a.ndim = ndim;
    for (int m = 0; m < ndim; m++) {
        a.idx[m] = idx[m];
        a.out[m] = (int32_t *)take(n * 4);
    }
    const dim3 grid = g_grid((int64_t)n);
    int st = DT_OK;
    auto finish = [&](int code) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(buf);
        return code;
    };
    // every index inside its mode
    int32_t *flag = keep;  // reused before the merge
    if (hipMemsetAsync(flag, 0, 4, ctx->stream) != hipSuccess) return finish(dt_fail(ctx, DT_EHIP, "merge: memset"));
    for (int m = 0; m < ndim; m++)
        hipLaunchKernelGGL(mg_check, grid, dim3(256), 0, ctx->stream, (const int32_t *)idx[m], nnz, (int64_t)dims[m], flag);
    int32_t hflag = 0;
    if (hipMemcpyAsync(&hflag, flag, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return finish(dt_fail(ctx, DT_EHIP, "dt_sp_merge_duplicates: index check failed"));
    if (hflag) return finish(dt_fail(ctx, DT_EINVAL, "dt_sp_merge_duplicates: an index lies outside its mode's range"));
    hipLaunchKernelGGL(mg_iota, grid, dim3(256), 0, ctx->stream, perm, nnz);
    for (int m = 0; m < ndim && st == DT_OK; m++) {
        hipLaunchKernelGGL(mg_keys, grid, dim3(256), 0, ctx->stream, (const int32_t *)idx[m], (const int32_t *)perm,
                           nnz, key);
        st = dt_radix_sort_u32(ctx, key, key2, perm, perm2, nnz, mg_bits((int64_t)dims[m]), stmp);
        std::swap(perm, perm2);
    }
    if (st != DT_OK) return finish(st);
    if (hipMemsetAsync(keep, 0, n * 4, ctx->stream) != hipSuccess)
        return finish(dt_fail(ctx, DT_EHIP, "dt_sp_merge_duplicates: memset failed"));
    hipLaunchKernelGGL(mg_heads, grid, dim3(256), 0, ctx->stream, a, (const int32_t *)perm, nnz, head);
    hipLaunchKernelGGL(mg_sums, grid, dim3(256), 0, ctx->stream, (const int32_t *)perm, (const int32_t *)head, nnz,
                       (const double *)val, keep, vsum);
    st = dt_exclusive_scan_i32(ctx, keep, pos, nnz, (int32_t *)stmp);
    if (st != DT_OK) return finish(st);
    hipLaunchKernelGGL(mg_compact, grid, dim3(256), 0, ctx->stream, a, (const int32_t *)keep, (const int32_t *)pos,
                       nnz, (const double *)vsum, vout);
    int32_t last[2];
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&last[0], pos + n - 1, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(&last[1], keep + n - 1, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return finish(dt_fail(ctx, DT_EHIP, "dt_sp_merge_duplicates: launch failed"));
    const int64_t nu = (int64_t)last[0] + last[1];
    for (int m = 0; m < ndim; m++)
        if (hipMemcpyAsync(idx[m], a.out[m], (size_t)nu * 4, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            return finish(dt_fail(ctx, DT_EHIP, "dt_sp_merge_duplicates: copy failed"));
    if (hipMemcpyAsync(val, vout, (size_t)nu * 8, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
        return finish(dt_fail(ctx, DT_EHIP, "dt_sp_merge_duplicates: copy failed"));
    *nnz_out = nu;
    return finish(DT_OK);
}
