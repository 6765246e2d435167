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
//   als_rows_kernel  : RB = 256 / R rows per workgroup (one output per thread):
//                      T = S[rows,:] Y (S symmetric, so a row of S is a
//                      contiguous column), out = T Rinv.  Optionally writes the
//                      FP32 / bf16 / transposed kernel views of the result.
//   als_gram_kernel  : partial Grams outᵀ out of 128-row blocks.
//   als_chol_kernel  : one workgroup: G = sum of the partials (fixed order),
//                      optional shift 11 (I R + R(R+1)) u tr(G) (Fukaya et al.
//                      2020), G = RᵀR, writes R^{-1}.
#include "dt_internal.hpp"
#include "dt_chol16.hpp"

#include <cmath>

constexpr int ALS_THREADS = 256;
constexpr int ALS_CH = 128;     // k-chunk staged in LDS
constexpr int ALS_MAXR = 64;
constexpr int ALS_GR = 64;      // rows per workgroup of the partial Gram

// rows per workgroup of als_rows_kernel: one output element per thread
template <int RMAX>
struct AlsRows {
    static constexpr int RB = RMAX <= 16 ? 16 : (RMAX <= 32 ? 8 : 4);
};

// mode 0: out = (S[rows,:] Y) Rinv      (S: I x I symmetric)
// mode 1: out = Y[rows,:] Rinv          (no S)
// Rinv == nullptr means identity.
template <int RMAX>
__global__ __launch_bounds__(ALS_THREADS) void als_rows_kernel(
    int mode, const double *__restrict__ S, const double *__restrict__ Y, int I, int R,
    const double *__restrict__ Rinv, double *__restrict__ out, float *__restrict__ U32,
    float *__restrict__ M32, uint16_t *__restrict__ Ub16) {
    constexpr int RB = AlsRows<RMAX>::RB;
    __shared__ double Ss[ALS_CH][RB + 1];     // S[t0+c, row0+r]
    __shared__ double Ys[ALS_CH][RMAX + 1];   // Y[t0+c, j]
    __shared__ double T[RB][RMAX + 1];        // rows x R
    __shared__ double Ri[RMAX][RMAX + 1];
    const int tid = threadIdx.x;
    const int row0 = blockIdx.x * RB;
    const int nrow = (I - row0) < RB ? (I - row0) : RB;

    for (int e = tid; e < R * R; e += ALS_THREADS) {
        const int i = e % R, j = e / R;
        Ri[i][j] = Rinv ? Rinv[e] : (i == j ? 1.0 : 0.0);
    }
    // outputs owned by this thread: (r, j) = (o % RB, o / RB), o = tid + q*256
    constexpr int NQ = (RB * RMAX + ALS_THREADS - 1) / ALS_THREADS;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) acc[q] = 0.0;

    if (mode == 0) {
        // chunks of ALS_CH rows of S and Y go through LDS; the next chunk's
        // global loads are issued before the current chunk's FMAs
        constexpr int NS = (ALS_CH * RB + ALS_THREADS - 1) / ALS_THREADS;
        constexpr int NY = (ALS_CH * RMAX + ALS_THREADS - 1) / ALS_THREADS;
        double sreg[NS], yreg[NY];
        auto fetch = [&](int t0) {
            const int ch = (I - t0) < ALS_CH ? (I - t0) : ALS_CH;
#pragma unroll
            for (int q = 0; q < NS; q++) {
                const int e = tid + ALS_THREADS * q;
                const int c = e % ALS_CH, r = e / ALS_CH;
                sreg[q] = (e < ALS_CH * RB && c < ch && r < nrow) ? S[(t0 + c) + (size_t)I * (row0 + r)] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < NY; q++) {
                const int e = tid + ALS_THREADS * q;
                const int c = e % ALS_CH, j = e / ALS_CH;
                yreg[q] = (j < R && c < ch) ? Y[(t0 + c) + (size_t)I * j] : 0.0;
            }
        };
        fetch(0);
        for (int t0 = 0; t0 < I; t0 += ALS_CH) {
#pragma unroll
            for (int q = 0; q < NS; q++) {
                const int e = tid + ALS_THREADS * q;
                if (e < ALS_CH * RB) Ss[e % ALS_CH][e / ALS_CH] = sreg[q];
            }
#pragma unroll
            for (int q = 0; q < NY; q++) {
                const int e = tid + ALS_THREADS * q;
                Ys[e % ALS_CH][e / ALS_CH] = yreg[q];
            }
            __syncthreads();
            if (t0 + ALS_CH < I) fetch(t0 + ALS_CH);
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int o = tid + q * ALS_THREADS;
                const int r = o % RB, j = o / RB;
                if (j < R) {
                    double s0 = acc[q], s1 = 0.0;
#pragma unroll 16
                    for (int c = 0; c < ALS_CH; c += 2) {
                        s0 = fma(Ss[c][r], Ys[c][j], s0);
                        s1 = fma(Ss[c + 1][r], Ys[c + 1][j], s1);
                    }
                    acc[q] = s0 + s1;
                }
            }
            __syncthreads();
        }
    } else {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int o = tid + q * ALS_THREADS;
            const int r = o % RB, j = o / RB;
            if (j < R && r < nrow) acc[q] = Y[(row0 + r) + (size_t)I * j];
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int o = tid + q * ALS_THREADS;
        const int r = o % RB, j = o / RB;
        if (j < R) T[r][j] = acc[q];
    }
    __syncthreads();
    // out = T Rinv (Rinv upper triangular or identity; full product is cheap)
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int o = tid + q * ALS_THREADS;
        const int r = o % RB, j = o / RB;
        if (j < R) {
            double s = 0.0;
            for (int m = 0; m < R; m++) s = fma(T[r][m], Ri[m][j], s);
            acc[q] = s;
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int o = tid + q * ALS_THREADS;
        const int r = o % RB, j = o / RB;
        if (j < R && r < nrow) {
            const size_t gi = (size_t)(row0 + r) + (size_t)I * j;
            out[gi] = acc[q];
            if (U32) U32[gi] = (float)acc[q];
            if (Ub16) Ub16[gi] = f32_to_bf16((float)acc[q]);
            if (M32) M32[j + (size_t)R * (row0 + r)] = (float)acc[q];
        }
    }
}

// part[w] = X[rows_w, :]ᵀ X[rows_w, :] for ALS_GR-row blocks w of X (I x R).
template <int RMAX>
__global__ __launch_bounds__(ALS_THREADS) void als_gram_kernel(const double *__restrict__ X, int I, int R,
                                                               double *__restrict__ part) {
    // partial Gram XᵀX of rows [rbeg, rend): the block is staged in LDS once
    // (coalesced loads; rows past rend and columns past R zero), then 16 x 16
    // tiles of the R x R result are dealt to the four waves in turn on FP64
    // MFMA: tile (ti, tj), step s, lane (c, g): Xs[4s + g][16 ti + c] as A and
    // Xs[4s + g][16 tj + c] as B; register v is entry (16 ti + 4v + g, 16 tj + c)
    constexpr int NT = RMAX / 16;
    __shared__ double Xs[ALS_GR][RMAX + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int rbeg = blockIdx.x * ALS_GR;
    const int rend = (rbeg + ALS_GR) < I ? (rbeg + ALS_GR) : I;
    for (int e = tid; e < ALS_GR * RMAX; e += ALS_THREADS) {
        const int r = e % ALS_GR, j = e / ALS_GR;
        Xs[r][j] = (rbeg + r < rend && j < R) ? X[(size_t)(rbeg + r) + (size_t)I * j] : 0.0;
    }
    __syncthreads();
    double *p = part + (size_t)blockIdx.x * R * R;
    for (int t = wv; t < NT * NT; t += 4) {
        const int ti = t % NT, tj = t / NT;
        doublex4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
        for (int s = 0; s < ALS_GR / 4; s++)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Xs[4 * s + g][16 * ti + c], Xs[4 * s + g][16 * tj + c], acc,
                                                       0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = 16 * ti + 4 * v + g, j = 16 * tj + c;
            if (i < R && j < R) p[i + (size_t)R * j] = acc[v];
        }
    }
}


// C += A B for 16 x 16 blocks on one wave's FP64 MFMA: a(k) gives A[c][k] and
// b(k) gives B[k][c] for the lane's column c = lane % 16 and k = 4 s + lane / 16;
// acc[v] ends as C[4 v + lane / 16][c].
template <typename FA, typename FB>
__device__ __forceinline__ doublex4 mm16(FA a, FB b, doublex4 acc) {
    const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int s = 0; s < 4; s++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a(4 * s + g), b(4 * s + g), acc, 0, 0, 0);
    return acc;
}

// Order 64 as a 4 x 4 grid of 16 x 16 blocks (G: LDS, row stride 65, upper
// part used), right-looking: for block column k, one wave factors the
// diagonal block (chol16_rinv on its symmetric copy: R_kk⁻¹), the waves form
// the block row R_kj = R_kk⁻ᵀ G_kj and update the trailing blocks
// G_ij -= R_kiᵀ R_kj (i <= j), one 16 x 16 tile per wave on FP64 MFMA.  Then
// R⁻¹ by block diagonals d = 1..3: X_ij = -R_ii⁻¹ sum_{k = i+1..j} R_ik X_kj,
// every X_kj of a smaller d being known (again one tile per wave).  Writes the
// R x R corner of R⁻¹ (column-major) unless a pivot fails, which ORs 4 into *flag.
__device__ void chol64_blocked(double (*G)[65], double (*Rb)[65], int R, double *__restrict__ Rinv_out,
                               int *__restrict__ flag) {
    __shared__ double Id[4][16][17];  // inverses of the diagonal blocks of R
    __shared__ double Xi[64][65];     // R^{-1}
    __shared__ double Tt[3][16][17];
    __shared__ double Gd[16][17];     // the diagonal block, both triangles
    __shared__ int bad_s;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), nwv = nt >> 6;
    const int lane = tid & 63, c = lane & 15, g = lane >> 4;
    const doublex4 zero4 = {0.0, 0.0, 0.0, 0.0};
    if (tid == 0) bad_s = 0;
    for (int kb = 0; kb < 4; kb++) {
        const int k0 = 16 * kb;
        __syncthreads();
        for (int e = tid; e < 256; e += nt) {
            const int i = e >> 4, k = e & 15;
            Gd[i][k] = i <= k ? G[k0 + i][k0 + k] : G[k0 + k][k0 + i];
        }
        __syncthreads();
        // wave 0 (the others return at once); a bad pivot ORs 4 into *flag
        // and is recorded for the early exit below
        chol16_rinv(Gd, Id[kb], 16, 0, 0, &bad_s);
        __syncthreads();
        const int ntl = 3 - kb;  // 16-column tiles right of the diagonal block
        for (int jt = wv; jt < ntl; jt += nwv) {
            const int j0 = k0 + 16 * (jt + 1);
            const doublex4 acc = mm16([&](int k) { return k <= c ? Id[kb][k][c] : 0.0; },  // (R_kk⁻ᵀ)[c][k]
                                      [&](int k) { return G[k0 + k][j0 + c]; }, zero4);
#pragma unroll
            for (int v = 0; v < 4; v++) Rb[k0 + 4 * v + g][j0 + c] = acc[v];
        }
        __syncthreads();
        for (int t = wv; t < ntl * (ntl + 1) / 2; t += nwv) {
            const int tj = t >= 3 ? 2 : (t >= 1 ? 1 : 0), ti = t - tj * (tj + 1) / 2;  // ti <= tj
            const int i0 = k0 + 16 * (ti + 1), j0 = k0 + 16 * (tj + 1);
            const doublex4 acc = mm16([&](int k) { return Rb[k0 + k][i0 + c]; },
                                      [&](int k) { return Rb[k0 + k][j0 + c]; }, zero4);
#pragma unroll
            for (int v = 0; v < 4; v++) G[i0 + 4 * v + g][j0 + c] -= acc[v];
        }
    }
    __syncthreads();
    if (bad_s) {
        if (tid == 0) atomicOr(flag, 4);
        return;
    }
    for (int e = tid; e < 64 * 64; e += nt) {
        const int i = e & 63, j = e >> 6;
        Xi[i][j] = ((i >> 4) == (j >> 4) && (i & 15) <= (j & 15)) ? Id[i >> 4][i & 15][j & 15] : 0.0;
    }
    for (int d = 1; d < 4; d++) {
        __syncthreads();
        for (int ib = wv; ib < 4 - d; ib += nwv) {
            const int jb = ib + d;
            doublex4 acc = zero4;
            for (int kb = ib + 1; kb <= jb; kb++)
                acc = mm16([&](int k) { return Rb[16 * ib + c][16 * kb + k]; },
                           [&](int k) { return Xi[16 * kb + k][16 * jb + c]; }, acc);
#pragma unroll
            for (int v = 0; v < 4; v++) Tt[ib][4 * v + g][c] = acc[v];
        }
        __syncthreads();
        for (int ib = wv; ib < 4 - d; ib += nwv) {
            const doublex4 acc = mm16([&](int k) { return k >= c ? Id[ib][c][k] : 0.0; },
                                      [&](int k) { return Tt[ib][k][c]; }, zero4);
#pragma unroll
            for (int v = 0; v < 4; v++) Xi[16 * ib + 4 * v + g][16 * (ib + d) + c] = -acc[v];
        }
    }
    __syncthreads();
    for (int e = tid; e < R * R; e += nt) Rinv_out[e] = Xi[e % R][e / R];
}

// One workgroup: G = sum_w part[w] (fixed order); optional shift; Rinv = chol(G)^{-1}.
// CM (16, 32 or 64 >= R) is the padded order of the register Cholesky; NTH
// threads (1024 at order 64: sixteen partials' loads in flight per pass and
// sixteen waves for the MFMA tiles of the blocked factorization).
template <int CM, int NTH>
__global__ __launch_bounds__(NTH) void als_chol_kernel(const double *__restrict__ part,
                                                               int nparts, int R, int I,
                                                               int shifted,
                                                               double *__restrict__ Rinv_out,
                                                               int *__restrict__ flag) {
    __shared__ double G[CM][CM + 1];
    __shared__ double X[CM][CM + 1];
    __shared__ double shift_s;
    const int tid = threadIdx.x;
    {
        // G = sum of the partials in a fixed order; each pass over w issues
        // this thread's loads for all its entries together
        constexpr int NE = (CM * CM + NTH - 1) / NTH;
        constexpr int PW = NE <= 4 ? 16 : (NE <= 8 ? 8 : 4);  // partials per pass
        const size_t RR = (size_t)R * R;
        double acc[NE];
#pragma unroll
        for (int q = 0; q < NE; q++) acc[q] = 0.0;
        // PW partials' loads in flight at a time, added in partial order
        // (a partial past nparts reads the last one and adds zero)
        for (int w0 = 0; w0 < nparts; w0 += PW) {
            double v[PW][NE];
#pragma unroll
            for (int u = 0; u < PW; u++) {
                const int w = w0 + u < nparts ? w0 + u : nparts - 1;
#pragma unroll
                for (int q = 0; q < NE; q++) {
                    const int e = tid + NTH * q;
                    v[u][q] = part[(size_t)w * RR + (e < R * R ? e : 0)];
                }
            }
#pragma unroll
            for (int u = 0; u < PW; u++)
#pragma unroll
                for (int q = 0; q < NE; q++) acc[q] += w0 + u < nparts ? v[u][q] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < NE; q++) {
            const int e = tid + NTH * q;
            if (e < R * R) G[e % R][e / R] = acc[q];
        }
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
    }
    __syncthreads();
    if (tid < R) G[tid][tid] += shift_s;
    // pad to CM with the identity: chol(diag(G, I)) = diag(chol(G), I)
    for (int e = tid; e < CM * CM; e += NTH) {
        const int i = e % CM, j = e / CM;
        if (i >= R || j >= R) G[i][j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    if constexpr (CM > 32) {
        chol64_blocked(G, X, R, Rinv_out, flag);
        return;
    }
    if (tid < 64) {
        // One wave, lane k < CM holding column k in registers (lanes past CM
        // repeat column CM - 1 and write nothing); every index below is a
        // compile-time constant after unrolling, so row j of the trailing
        // update comes from v_readlane (lane i's col[j] = G[j][i]).
        const int k = tid < CM ? tid : CM - 1;
        double col[CM];
#pragma unroll
        for (int i = 0; i < CM; i++) col[i] = G[i][k];
        int bad = 0;
        double invr[CM];  // 1 / R[j][j]: one division per pivot, none in the back substitution
#pragma unroll
        for (int j = 0; j < CM; j++) {
            double d = rl_f64(col[j], j);
            if (!(d > 0.0) || !isfinite(d)) {
                if (!bad) bad = j + 1;
                d = 1.0;
            }
            const double rjj = sqrt(d), ir = 1.0 / rjj;
            invr[j] = ir;
            const double gjk = col[j];
            const double f = gjk * (ir * ir);
#pragma unroll
            for (int i = j + 1; i < CM; i++) col[i] -= rl_f64(col[j], i) * f;
            col[j] = (k == j) ? rjj : gjk * ir;  // row j of R (meaningful for k >= j)
        }
        if (tid < CM)
#pragma unroll
            for (int i = 0; i < CM; i++) X[i][k] = col[i];  // X = R (upper part)
        if (bad && tid == 0) atomicOr(flag, 4);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // column k of R^{-1}: x[i] = (delta_ik - sum_{m>i} R[i][m] x[m]) / R[i][i];
        // x[m] = 0 for m > k falls out of the recurrence
        double x[CM];
#pragma unroll
        for (int i = CM - 1; i >= 0; i--) {
            double s0 = (i == k) ? 1.0 : 0.0, s1 = 0.0;
#pragma unroll
            for (int m = i + 1; m < CM; m += 2) {
                s0 -= X[i][m] * x[m];
                if (m + 1 < CM) s1 -= X[i][m + 1] * x[m + 1];
            }
            x[i] = (s0 + s1) * invr[i];
        }
        if (!bad && tid < R)
#pragma unroll
            for (int i = 0; i < CM; i++)
                if (i < R) Rinv_out[i + (size_t)R * k] = x[i];
    }
}

static int launch_chol(dt_ctx *ctx, const double *part, int nparts, int R, int I, int shifted, double *ri,
                       int *flag) {
    if (R <= 16) {
        DT_LAUNCH("als_chol", (als_chol_kernel<16, ALS_THREADS>), dim3(1), dim3(ALS_THREADS), 0, part, nparts, R, I, shifted, ri,
                  flag);
    } else if (R <= 32) {
        DT_LAUNCH("als_chol", (als_chol_kernel<32, ALS_THREADS>), dim3(1), dim3(ALS_THREADS), 0, part, nparts, R, I, shifted, ri,
                  flag);
    } else {
        DT_LAUNCH("als_chol", (als_chol_kernel<64, 1024>), dim3(1), dim3(1024), 0, part, nparts, R, I, shifted, ri,
                  flag);
    }
    return DT_OK;
}

// ---- the whole update in one workgroup (I * CM <= AF_MAXIR) -------------------
// Y and W (I x CM, row-major, columns R..CM-1 zero) live in LDS for all the
// steps; W = S Q runs on v_mfma_f64_16x16x4f64 (wave w takes 16-row tiles
// w, w + 16, ...; lane (c, g) holds S[row c][k + g] as A and Q[k + g][c] as B;
// output register v of lane (c, g) is row 4v + g, column c), the Gram, the
// register Cholesky and the triangular products run on the VALU.  The same
// steps as the multi-workgroup path:
//   Q_0 = U;  for k < inner: W = S Q_k, R_k = chol(WᵀW + shift), Q_{k+1} = W R_k^{-1};
//   then R = chol(QᵀQ), U = Q R^{-1}.
constexpr int AF_THREADS = 512, AF_MAXIR = 8192;


template <int CM>
__device__ __forceinline__ void af_chol(double (*G)[CM + 1], double (*Ri)[CM + 1], double (*X)[CM + 1], int R,
                                        int I, int shifted, int *flag) {
    // G (R x R, LDS) -> Ri = chol(G + shift)^{-1} (CM x CM, identity-padded); one wave
    if constexpr (CM == 16) {
        chol16_rinv(G, Ri, R, I, shifted, flag);
        return;
    }
    const int tid = threadIdx.x;
    if (tid < 64) {
        double shift = 0.0;
        if (shifted) {
            double tr = 0.0;
            for (int i = 0; i < R; i++) tr += G[i][i];
            shift = 11.0 * ((double)I * R + (double)R * (R + 1)) * 1.1102230246251565e-16 * tr;
        }
        const int k = tid < CM ? tid : CM - 1;
        double col[CM];
#pragma unroll
        for (int i = 0; i < CM; i++)
            col[i] = (i < R && k < R) ? G[i][k] + (i == k ? shift : 0.0) : (i == k ? 1.0 : 0.0);
        int bad = 0;
        // CM <= 16: one division per pivot, 1 / R[j][j] kept for the back
        // substitution; wider panels keep the two-division form (registers)
        double invr[CM <= 16 ? CM : 1];
#pragma unroll
        for (int j = 0; j < CM; j++) {
            double d = rl_f64(col[j], j);
            if (!(d > 0.0) || !isfinite(d)) {
                if (!bad) bad = j + 1;
                d = 1.0;
            }
            const double gjk = col[j];
            if constexpr (CM <= 16) {
                const double rjj = sqrt(d), ir = 1.0 / rjj;
                invr[j] = ir;
                const double f = gjk * (ir * ir);
#pragma unroll
                for (int i = j + 1; i < CM; i++) col[i] -= rl_f64(col[j], i) * f;
                col[j] = (k == j) ? rjj : gjk * ir;
            } else {
                const double rjj = sqrt(d), inv_d = 1.0 / d;
#pragma unroll
                for (int i = j + 1; i < CM; i++) col[i] -= rl_f64(col[j], i) * gjk * inv_d;
                col[j] = (k == j) ? rjj : gjk / rjj;
            }
        }
        if (tid < CM)
#pragma unroll
            for (int i = 0; i < CM; i++) X[i][k] = col[i];
        if (bad && tid == 0) atomicOr(flag, 4);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double x[CM];
#pragma unroll
        for (int i = CM - 1; i >= 0; i--) {
            double s0 = (i == k) ? 1.0 : 0.0, s1 = 0.0;
#pragma unroll
            for (int m = i + 1; m < CM; m += 2) {
                s0 -= X[i][m] * x[m];
                if (m + 1 < CM) s1 -= X[i][m + 1] * x[m + 1];
            }
            if constexpr (CM <= 16)
                x[i] = (s0 + s1) * invr[i];
            else
                x[i] = (s0 + s1) / X[i][i];
        }
        if (tid < CM)
#pragma unroll
            for (int i = 0; i < CM; i++) Ri[i][k] = (i < R && k < R) ? x[i] : (i == k ? 1.0 : 0.0);
    }
}

typedef double doublex4 __attribute__((ext_vector_type(4)));

template <int CM>
__global__ __launch_bounds__(AF_THREADS) void als_fused_kernel(const double *__restrict__ S, double *__restrict__ U,
                                                               int I, int R, int inner, int *__restrict__ flag,
                                                               float *__restrict__ U32, float *__restrict__ M32,
                                                               uint16_t *__restrict__ Ub16) {
    constexpr int NW = AF_THREADS / 64, NCT = CM / 16;
    static_assert(NW * NCT * NCT * 256 <= AF_MAXIR, "Gram tiles must fit in a panel");
    __shared__ double Y[AF_MAXIR], W[AF_MAXIR];
    __shared__ double G[CM][CM + 1], Ri[CM][CM + 1], X[CM][CM + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int n = I * CM, ntile = (I + 15) / 16;
    for (int e = tid; e < n; e += AF_THREADS) {
        const int i = e / CM, r = e % CM;
        Y[e] = r < R ? U[i + (size_t)I * r] : 0.0;
    }
    for (int e = tid; e < CM * CM; e += AF_THREADS) Ri[e / CM][e % CM] = (e / CM == e % CM) ? 1.0 : 0.0;
    __syncthreads();
    // dst = src Ri: per 16-row tile and column tile, K = CM on MFMA
    auto times_ri = [&](const double *src, double *dst) {
        for (int t = wv; t < ntile * NCT; t += NW) {
            const int rt = t % ntile, ct = t / ntile;
            const int row = 16 * rt + c;
            doublex4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k0 = 0; k0 < CM; k0 += 4) {
                const double a = row < I ? src[row * CM + k0 + g] : 0.0;
                d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Ri[k0 + g][16 * ct + c], d, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = 16 * rt + 4 * v + g;
                if (i < I) dst[i * CM + 16 * ct + c] = d[v];
            }
        }
    };
    // G = srcᵀ src (R x R): every wave sums its own rows on MFMA, then the waves
    // are added in order; the per-wave tiles go to `scratch` (the other panel,
    // NW * NCT^2 * 256 <= AF_MAXIR doubles)
    auto gram = [&](const double *src, double *scratch) {
        double(*Gp)[NCT * NCT][4][64] = (double(*)[NCT * NCT][4][64])scratch;
        doublex4 d[NCT * NCT];
#pragma unroll
        for (int q = 0; q < NCT * NCT; q++) d[q] = doublex4{0.0, 0.0, 0.0, 0.0};
        for (int k0 = 4 * wv; k0 < I; k0 += 4 * NW) {
            const int kk = k0 + g;
            double v[NCT];
#pragma unroll
            for (int h = 0; h < NCT; h++) v[h] = kk < I ? src[kk * CM + 16 * h + c] : 0.0;
#pragma unroll
            for (int a = 0; a < NCT; a++)
#pragma unroll
                for (int b = 0; b < NCT; b++)
                    d[a * NCT + b] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[a], v[b], d[a * NCT + b], 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < NCT * NCT; q++)
#pragma unroll
            for (int v = 0; v < 4; v++) Gp[wv][q][v][lane] = d[q][v];
        __syncthreads();
        for (int e = tid; e < NCT * NCT * 256; e += AF_THREADS) {
            const int q = e / 256, v = (e / 64) % 4, l = e % 64;
            double sum = 0.0;
            for (int w = 0; w < NW; w++) sum += Gp[w][q][v][l];
            // lane l = (col, grp), register v: tile row 4v + grp, column col
            const int a = (q / NCT) * 16 + 4 * v + (l >> 4), b = (q % NCT) * 16 + (l & 15);
            if (a < R && b < R) G[a][b] = sum;
        }
    };
    for (int k = 0;; k++) {
        times_ri(Y, W);  // Q_k = Y R^{-1}
        __syncthreads();
        if (k == inner) break;
        // Y = S Q_k on MFMA, 16 steps of S loads in flight
        for (int t = wv; t < ntile * NCT; t += NW) {
            const int rt = t % ntile, ct = t / ntile;
            const int row = 16 * rt + c;
            const int rowc = row < I ? row : I - 1;
            doublex4 d = {0.0, 0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < I; k0 += 64) {
                double a[16];
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const int kk = k0 + 4 * j + g;
                    a[j] = S[(size_t)rowc + (size_t)I * (kk < I ? kk : I - 1)];
                }
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const int kk = k0 + 4 * j + g;
                    const bool ok = row < I && kk < I;
                    const double b = kk < I ? W[kk * CM + 16 * ct + c] : 0.0;
                    d = __builtin_amdgcn_mfma_f64_16x16x4f64(ok ? a[j] : 0.0, b, d, 0, 0, 0);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = 16 * rt + 4 * v + g;
                if (i < I) Y[i * CM + 16 * ct + c] = d[v];
            }
        }
        __syncthreads();
        gram(Y, W);
        __syncthreads();
        af_chol<CM>(G, Ri, X, R, I, 1, flag);
        __syncthreads();
    }
    // W = Q_inner; R = chol(QᵀQ); U = Q R^{-1}
    gram(W, Y);
    __syncthreads();
    af_chol<CM>(G, Ri, X, R, I, 0, flag);
    __syncthreads();
    times_ri(W, Y);
    __syncthreads();
    for (int e = tid; e < I * R; e += AF_THREADS) {
        const int i = e % I, r = e / I;
        const double v = Y[i * CM + r];
        U[e] = v;
        if (U32) U32[e] = (float)v;
        if (Ub16) Ub16[e] = f32_to_bf16((float)v);
        if (M32) M32[r + (size_t)R * i] = (float)v;
    }
}

// ---- the update over row tiles: one launch per step --------------------------
// For R <= 16 and 160 <= I <= 512 the iteration runs on ceil(I / 16)
// workgroups, one per 16-row tile, and each step is one launch (the launch
// boundary is the only grid-wide synchronisation).  The steps are the fused
// kernel's, with the product grouped as Y_{k+1} = (S Y_k) R_k^{-1}:
//   AT_FIRST: Y_1 = S U
//   AT_STEP : R_k = chol(Y_kᵀ Y_k + shift), Y_{k+1} = (S Y_k) R_k^{-1}
//   AT_Q    : R = chol(Y_innerᵀ Y_inner + shift), Q = Y_inner R^{-1}
//   AT_LAST : R = chol(QᵀQ), U = Q R^{-1} and its kernel views
// A workgroup writes its 16 rows of the new panel (I x 16 row-major FP64,
// columns R..15 zero) and their 16 x 16 partial Gram; the next launch sums
// the partials in tile order and every workgroup factors the same Gram
// (af_chol on wave 0) while its S rows and the panel are in flight.
// S Y_k runs on waves 1..4, beside the Cholesky: wave 1 + w takes the 64-wide
// k chunks w and w + 4, lane (c, g) holding S[row c][k0 + 4s + g] (a column of
// the symmetric S) and Y_k[k0 + 4s + g][c] for the 16 MFMA steps s; the four
// waves' 16 x 16 partials are added in wave order.  Every launch reads S's rows of its tile (32 I bytes), the panel and
// the partials: a few hundred KB per launch, latency-bound.
enum { AT_FIRST = 0, AT_STEP = 1, AT_Q = 2, AT_LAST = 3 };
constexpr int AT_MINI = 160, AT_MAXI = 512;

constexpr int AT_THREADS = 320;  // wave 0: Cholesky; waves 1..4: S Y_k

template <int MODE>
__global__ __launch_bounds__(AT_THREADS) void als_tile(const double *__restrict__ S, const double *__restrict__ Yin,
                                                const double *__restrict__ Pin, double *__restrict__ Yout,
                                                double *__restrict__ Pout, int I, int R, int *__restrict__ flag,
                                                double *__restrict__ U, float *__restrict__ U32,
                                                float *__restrict__ M32, uint16_t *__restrict__ Ub16) {
    constexpr int CM = 16;
    constexpr bool HAS_S = MODE == AT_FIRST || MODE == AT_STEP;
    __shared__ double G[CM][CM + 1], Ri[CM][CM + 1], X[CM][CM + 1];
    __shared__ double Zp[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, mw = wv - 1;
    const int c = lane & 15, g = lane >> 4;
    const int T = (I + 15) / 16, r0 = blockIdx.x * 16;
    const bool ent = tid < 256;  // one entry of a 16 x 16 tile per thread
    const int row = r0 + c, rowc = row < I ? row : I - 1;
    // the previous launch's partial Grams first (they gate the Cholesky):
    // all T <= 32 loads in flight together, summed below in tile order
    constexpr int TMAX = AT_MAXI / 16;
    double pv[TMAX];
    if (MODE != AT_FIRST && ent) {
#pragma unroll
        for (int t = 0; t < TMAX; t++) pv[t] = t < T ? Pin[(size_t)t * 256 + tid] : 0.0;
    }
    if (!HAS_S && ent) {
        // AT_Q / AT_LAST: the tile's rows of the panel, [row][column]
        const int rr = r0 + (tid >> 4);
        Zp[0][tid] = rr < I ? Yin[(size_t)rr * CM + (tid & 15)] : 0.0;
    }
    // the operands of S Y_k (indices clamped into range, values outside it
    // zeroed).  Loads complete in issue order, so after a Cholesky they are
    // issued only once the partial Grams are in: then they overlap the
    // factorization instead of delaying its start.
    double a[2][16], b[2][16];
    auto load_sy = [&]() {
        if (!HAS_S || wv == 0) return;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int k0 = (mw + 4 * h) * 64;
            if (k0 >= I) continue;
#pragma unroll
            for (int s = 0; s < 16; s++) {
                const int kk = k0 + 4 * s + g;
                const int kc = kk < I ? kk : I - 1;
                const double sv = S[(size_t)kc + (size_t)I * rowc];
                const double yv = MODE == AT_FIRST ? U[(size_t)kc + (size_t)I * (c < R ? c : 0)] : Yin[(size_t)kc * CM + c];
                a[h][s] = kk < I && row < I ? sv : 0.0;
                b[h][s] = kk < I && (MODE != AT_FIRST || c < R) ? yv : 0.0;
            }
        }
    };
    if (MODE == AT_FIRST) load_sy();
    if (MODE != AT_FIRST) {
        if (ent) {
            double sum = 0.0;
#pragma unroll
            for (int t = 0; t < TMAX; t++)
                if (t < T) sum += pv[t];
            const int ga = tid >> 4, gb = tid & 15;
            if (ga < R && gb < R) G[ga][gb] = sum;
        }
        __syncthreads();
        load_sy();
        af_chol<CM>(G, Ri, X, R, I, MODE == AT_LAST ? 0 : 1, flag);
    } else if (ent) {
        Ri[tid >> 4][tid & 15] = (tid >> 4) == (tid & 15) ? 1.0 : 0.0;
    }
    if (HAS_S && wv > 0) {
        doublex4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int h = 0; h < 2; h++) {
            if ((mw + 4 * h) * 64 >= I) break;
#pragma unroll
            for (int s = 0; s < 16; s++) d = __builtin_amdgcn_mfma_f64_16x16x4f64(a[h][s], b[h][s], d, 0, 0, 0);
        }
        // register v of lane (c, g) holds row 4v + g, column c
#pragma unroll
        for (int v = 0; v < 4; v++) Zp[mw][(4 * v + g) * 16 + c] = d[v];
    }
    __syncthreads();
    if (HAS_S) {
        const double z = ent ? ((Zp[0][tid] + Zp[1][tid]) + Zp[2][tid]) + Zp[3][tid] : 0.0;
        __syncthreads();
        if (ent) Zp[0][tid] = z;
        __syncthreads();
    }
    if (wv != 0) return;
    // the tile's rows of the new panel: Z R^{-1} (K = 16 on the MFMA)
    doublex4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; s++)
        y = __builtin_amdgcn_mfma_f64_16x16x4f64(Zp[0][c * 16 + 4 * s + g], Ri[4 * s + g][c], y, 0, 0, 0);
    if (MODE == AT_LAST) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = r0 + 4 * v + g;
            if (i < I && c < R) {
                const double val = y[v];
                U[(size_t)i + (size_t)I * c] = val;
                if (U32) U32[(size_t)i + (size_t)I * c] = (float)val;
                if (Ub16) Ub16[(size_t)i + (size_t)I * c] = f32_to_bf16((float)val);
                if (M32) M32[c + (size_t)R * i] = (float)val;
            }
        }
        return;
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int i = r0 + 4 * v + g;
        Zp[1][(4 * v + g) * 16 + c] = i < I ? y[v] : 0.0;
        if (i < I) Yout[(size_t)i * CM + c] = y[v];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // partial Gram of the 16 rows: A[i][k] = Y[k][i] and B[k][j] = Y[k][j]
    // are the same value in lane (c, g) at step s (k = 4s + g)
    doublex4 p = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const double yy = Zp[1][(4 * s + g) * 16 + c];
        p = __builtin_amdgcn_mfma_f64_16x16x4f64(yy, yy, p, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) Pout[(size_t)blockIdx.x * 256 + (4 * v + g) * 16 + c] = p[v];
}

// ws: two panels (I x 16) and two sets of ceil(I / 16) partial Grams (16 x 16).
static int als_update_tiles(dt_ctx *ctx, const double *S, double *U, int I, int R, int inner, double *ws,
                            int *flag, float *U32, float *M32, uint16_t *Ub16) {
    const int T = (I + 15) / 16;
    double *Y[2] = {ws, ws + (size_t)I * 16};
    double *P[2] = {ws + (size_t)2 * I * 16, ws + (size_t)2 * I * 16 + (size_t)T * 256};
    DT_LAUNCH("als_tile", als_tile<AT_FIRST>, dim3(T), dim3(AT_THREADS), 0, S, (const double *)nullptr,
              (const double *)nullptr, Y[0], P[0], I, R, flag, U, (float *)nullptr, (float *)nullptr,
              (uint16_t *)nullptr);
    int cur = 0;
    for (int k = 1; k < inner; k++, cur ^= 1)
        DT_LAUNCH("als_tile", als_tile<AT_STEP>, dim3(T), dim3(AT_THREADS), 0, S, (const double *)Y[cur],
                  (const double *)P[cur], Y[cur ^ 1], P[cur ^ 1], I, R, flag, (double *)nullptr, (float *)nullptr,
                  (float *)nullptr, (uint16_t *)nullptr);
    DT_LAUNCH("als_tile", als_tile<AT_Q>, dim3(T), dim3(AT_THREADS), 0, S, (const double *)Y[cur], (const double *)P[cur],
              Y[cur ^ 1], P[cur ^ 1], I, R, flag, (double *)nullptr, (float *)nullptr, (float *)nullptr,
              (uint16_t *)nullptr);
    cur ^= 1;
    DT_LAUNCH("als_tile", als_tile<AT_LAST>, dim3(T), dim3(AT_THREADS), 0, S, (const double *)Y[cur],
              (const double *)P[cur], (double *)nullptr, (double *)nullptr, I, R, flag, U, U32, M32, Ub16);
    return DT_OK;
}

// Mode 0 on FP64 MFMA for I % 128 == 0: out = (S Y) Rinv, a workgroup owning
// 16 rows.  Wave w sums k over [w I/4, (w + 1) I/4) in chunks of 16: lane
// (c, g) loads S[row0 + c][k0 + 4g .. +3] (S symmetric, so that is 32
// contiguous bytes of column row0 + c) and Y[k0 + 4g .. +3][16t + c] for the
// four 16-column tiles t; MFMA step s pairs k = k0 + 4g + s in both operands,
// and register v of accumulator t ends as (S Y)[row0 + 4v + g][16t + c].  Two
// register sets of chunks alternate.  The waves' partials are added in wave
// order through LDS, then multiplied by Rinv (identity when null).
__global__ __launch_bounds__(ALS_THREADS) void als_sy_mfma(const double *__restrict__ S, const double *__restrict__ Y,
                                                          int I, int R, const double *__restrict__ Rinv,
                                                          double *__restrict__ out) {
    __shared__ double P[16][65];
    __shared__ double Ri[64][65];
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 15, g = lane >> 4;
    const int row0 = blockIdx.x * 16;
    for (int e = tid; e < 64 * 64; e += ALS_THREADS) {
        const int i = e & 63, j = e >> 6;
        Ri[i][j] = (i < R && j < R) ? (Rinv ? Rinv[i + (size_t)R * j] : (i == j ? 1.0 : 0.0)) : 0.0;
    }
    const int kw = I / 4, kbeg = wv * kw, kend = kbeg + kw;
    const double *sp = S + (size_t)(row0 + c) * I + 4 * g;
    const double *yp[4];
    bool yin[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int col = 16 * t + c;
        yin[t] = col < R;
        yp[t] = Y + (size_t)(yin[t] ? col : R - 1) * I + 4 * g;
    }
    doublex4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = doublex4{0.0, 0.0, 0.0, 0.0};
    auto ld = [&](int k0, doublex4 &a, doublex4 (&b)[4]) {
        a = *(const doublex4 *)(sp + k0);
#pragma unroll
        for (int t = 0; t < 4; t++) b[t] = *(const doublex4 *)(yp[t] + k0);
    };
    auto mma = [&](const doublex4 &a, const doublex4 (&b)[4]) {
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int t = 0; t < 4; t++)
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], yin[t] ? b[t][s] : 0.0, acc[t], 0, 0, 0);
    };
    doublex4 aX, bX[4], aY, bY[4];
    ld(kbeg, aX, bX);
    ld(kbeg + 16, aY, bY);
    for (int k0 = kbeg; k0 < kend; k0 += 32) {
        // past the range the last chunk is loaded again and not used
        mma(aX, bX);
        ld(k0 + 32 < kend ? k0 + 32 : k0, aX, bX);
        mma(aY, bY);
        ld(k0 + 48 < kend ? k0 + 48 : k0 + 16, aY, bY);
    }
    for (int w = 0; w < 4; w++) {
        if (wv == w) {
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const double x = acc[t][v];
                    P[4 * v + g][16 * t + c] = w > 0 ? P[4 * v + g][16 * t + c] + x : x;
                }
        }
        __syncthreads();
    }
    for (int e = tid; e < 16 * R; e += ALS_THREADS) {
        const int i = e & 15, j = e >> 4;
        double s0 = 0.0, s1 = 0.0;
        for (int m = 0; m < 64; m += 2) {
            s0 = fma(P[i][m], Ri[m][j], s0);
            s1 = fma(P[i][m + 1], Ri[m + 1][j], s1);
        }
        out[(size_t)(row0 + i) + (size_t)I * j] = s0 + s1;
    }
}

template <int RMAX>
static int launch_rows(dt_ctx *ctx, int mode, const double *S, const double *Y, int I, int R,
                       const double *Rinv, double *out, double *part, float *U32, float *M32,
                       uint16_t *Ub16) {
    if (mode == 0 && R > 32 && I >= 128 && I % 128 == 0 && out != Y && !U32 && !M32 && !Ub16) {
        DT_LAUNCH("als_rows", als_sy_mfma, dim3(I / 16), dim3(ALS_THREADS), 0, S, Y, I, R, Rinv, out);
    } else {
        const int nb = (I + AlsRows<RMAX>::RB - 1) / AlsRows<RMAX>::RB;
        DT_LAUNCH("als_rows", als_rows_kernel<RMAX>, dim3(nb), dim3(ALS_THREADS), 0, mode, S, Y, I, R, Rinv,
                  out, U32, M32, Ub16);
    }
    if (part) {
        const int ng = (I + ALS_GR - 1) / ALS_GR;
        DT_LAUNCH("als_gram", als_gram_kernel<RMAX>, dim3(ng), dim3(ALS_THREADS), 0, out, I, R, part);
    }
    return DT_OK;
}

static int rows(dt_ctx *ctx, int mode, const double *S, const double *Y, int I, int R,
                const double *Rinv, double *out, double *part, float *U32 = nullptr,
                float *M32 = nullptr, uint16_t *Ub16 = nullptr) {
    if (R <= 16) return launch_rows<16>(ctx, mode, S, Y, I, R, Rinv, out, part, U32, M32, Ub16);
    if (R <= 32) return launch_rows<32>(ctx, mode, S, Y, I, R, Rinv, out, part, U32, M32, Ub16);
    return launch_rows<64>(ctx, mode, S, Y, I, R, Rinv, out, part, U32, M32, Ub16);
}

// Rank check of the Q-only Cholesky-QR3: diag(R) of X = Q R is the product of
// the three passes' diagonals 1 / Rinv_k[j][j].  d[j] accumulates it; after the
// last pass a column whose diagonal is at most 10 R u ||X||_F (||X||_F^2 = the
// trace of the first Gram, in d[R]) means X is rank-deficient to working
// precision: ORs 4 into *flag.  One wave, lane j for column j (R <= 64).
__global__ void cholqr_rank_track(const double *__restrict__ Rinv, const double *__restrict__ G0, int R, int pass,
                                  double *__restrict__ d, int *__restrict__ flag) {
    const int j = threadIdx.x;
    if (blockIdx.x != 0 || j >= 64) return;  // one full wave: every lane below takes part
    double tr;
    if (pass == 0) {
        tr = j < R ? G0[j + (size_t)R * j] : 0.0;
        for (int o = 32; o > 0; o >>= 1) tr += __shfl_xor(tr, o);
        if (j == 0) d[R] = tr;
    } else {
        tr = d[R];
    }
    bool bad = false;
    if (j < R) {
        const double rj = 1.0 / Rinv[j + (size_t)R * j];
        const double dj = pass == 0 ? rj : d[j] * rj;
        d[j] = dj;
        bad = pass == 2 && !(dj > 10.0 * R * 1.1102230246251565e-16 * sqrt(tr));
    }
    if (__ballot(bad) != 0 && j == 0) atomicOr(flag, 4);
}

// Q <- the orthonormal factor of X (I x R FP64, R <= 64) with positive R
// diagonal: shifted Cholesky-QR, then two plain passes (Fukaya et al. 2020),
// each pass one als_chol (partial Grams summed, R^{-1}) and one als_rows
// (X <- X R^{-1} in place) that also leaves the next pass's partial Grams.
// Nothing waits for the host; a Cholesky breakdown ORs 4 into *flag, and with
// rank_check so does a rank-deficient X (cholqr_rank_track).
// ws: dt_cholqr3_ws_doubles(I, R) doubles.
size_t dt_cholqr3_ws_doubles(size_t I, size_t R) {
    return ((I + ALS_GR - 1) / ALS_GR + 2) * R * R + R + 1;
}

int dt_cholqr3_rows(dt_ctx *ctx, double *X, int I, int R, double *ws, int *flag, bool rank_check) {
    if (R < 1 || R > ALS_MAXR) return dt_fail(ctx, DT_EUNSUPPORTED, "Cholesky-QR: %d columns > %d", R, ALS_MAXR);
    const int nb = (I + ALS_GR - 1) / ALS_GR;
    double *part = ws, *ri = ws + (size_t)nb * R * R;
    if (R <= 16) {
        DT_LAUNCH("als_gram", als_gram_kernel<16>, dim3(nb), dim3(ALS_THREADS), 0, X, I, R, part);
    } else if (R <= 32) {
        DT_LAUNCH("als_gram", als_gram_kernel<32>, dim3(nb), dim3(ALS_THREADS), 0, X, I, R, part);
    } else {
        DT_LAUNCH("als_gram", als_gram_kernel<64>, dim3(nb), dim3(ALS_THREADS), 0, X, I, R, part);
    }
    double *G0 = ri + (size_t)R * R, *dtrack = G0 + (size_t)R * R;
    if (rank_check) DT_TRY(dt_part_reduce(ctx, part, nb, R * R, G0));
    for (int pass = 0; pass < 3; pass++) {
        DT_TRY(launch_chol(ctx, part, nb, R, I, pass == 0 ? 1 : 0, ri, flag));
        if (rank_check)
            DT_LAUNCH("cholqr_small", cholqr_rank_track, dim3(1), dim3(64), 0, (const double *)ri, (const double *)G0, R,
                      pass, dtrack, flag);
        DT_TRY(rows(ctx, 1, nullptr, X, I, R, ri, X, pass < 2 ? part : nullptr));
    }
    return DT_OK;
}

// Workspace needs: part (ceil(I/128) R^2) + 2 Rinv (R^2) + 2 panels (I R)
// doubles; ranks above ALS_MAXR use a Gram, an R^{-1}, a reciprocal diagonal
// and the two panels; the row-tile update two I x 16 panels and two sets of
// ceil(I / 16) 16 x 16 partial Grams.
size_t dt_als_ws_doubles(size_t I, size_t R) {
    const size_t rows = ((I + ALS_GR - 1) / ALS_GR) * R * R + 2 * R * R + 2 * I * R + 64;
    const size_t any = 2 * R * R + R + 2 * I * R + 64;
    const size_t tiles = 2 * I * 16 + 2 * ((I + 15) / 16) * 256 + 64;
    const size_t m = rows > any ? rows : any;
    return m > tiles ? m : tiles;
}

// Ranks above ALS_MAXR: the same iteration with every I x R product on the
// tiled FP64 GEMM and the R x R Cholesky in global memory (one workgroup):
//   Q_0 = U;  for k < inner: Y = S Q_k, G = YᵀY, Q_{k+1} = Y chol(G + shift)^{-1};
//   then U = Q chol(QᵀQ)^{-1}.
static int als_update_any(dt_ctx *ctx, const double *S, double *U, int I, int R, int inner, double *ws,
                          int *flag, float *U32, float *M32, uint16_t *Ub16) {
    const size_t RR = (size_t)R * R, IR = (size_t)I * R;
    double *G = ws, *Ri = G + RR, *invd = Ri + RR, *Y = invd + R, *Q = Y + IR;
    const double *cur = U;
    for (int k = 0; k < inner; k++) {
        DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, I, R, I, 1.0, S, I, cur, I, 0.0, Y, I));
        DT_TRY(dt_gemm_f64_internal(ctx, 1, 0, R, R, I, 1.0, Y, I, Y, I, 0.0, G, R));
        DT_TRY(dt_chol_rinv_ws(ctx, G, R, true, (double)I, Ri, invd, flag));
        DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, I, R, R, 1.0, Y, I, Ri, R, 0.0, Q, I));
        cur = Q;
    }
    DT_TRY(dt_gemm_f64_internal(ctx, 1, 0, R, R, I, 1.0, cur, I, cur, I, 0.0, G, R));
    DT_TRY(dt_chol_rinv_ws(ctx, G, R, false, (double)I, Ri, invd, flag));
    DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, I, R, R, 1.0, cur, I, Ri, R, 0.0, Y, I));
    DT_TRY(dt_copy_f64(ctx, U, Y, IR));
    return dt_f64_to_views(ctx, U, (size_t)I, (size_t)R, U32, M32, Ub16, nullptr);
}

// U (I x R, orthonormal, FP64) <- orthonormal basis of span(S^inner U).
// U32 (I x R), M32 (R x I), Ub16 (I x R) receive the kernel views (may be null).
int dt_als_update(dt_ctx *ctx, const double *S, double *U, int I, int R, int inner, double *ws,
                  int *flag, float *U32, float *M32, uint16_t *Ub16) {
    if (R < 1 || R > I) return dt_fail(ctx, DT_EINVAL, "ALS update: rank %d not in [1, %d]", R, I);
    if (R > ALS_MAXR) return als_update_any(ctx, S, U, I, R, inner, ws, flag, U32, M32, Ub16);
    if (R <= 16 && I >= AT_MINI && I <= AT_MAXI && inner >= 1)
        return als_update_tiles(ctx, S, U, I, R, inner, ws, flag, U32, M32, Ub16);
    if (R <= 16 && I * 16 <= AF_MAXIR) {
        DT_LAUNCH("als_fused", als_fused_kernel<16>, dim3(1), dim3(AF_THREADS), 0, S, U, I, R, inner, flag, U32, M32,
                  Ub16);
        return DT_OK;
    }
    if (R <= 32 && I * 32 <= AF_MAXIR) {
        DT_LAUNCH("als_fused", als_fused_kernel<32>, dim3(1), dim3(AF_THREADS), 0, S, U, I, R, inner, flag, U32, M32,
                  Ub16);
        return DT_OK;
    }
    const int nb = (I + ALS_GR - 1) / ALS_GR;  // partial Grams
    double *part = ws;
    double *Ri = part + (size_t)nb * R * R;
    double *Yb[2] = {Ri + 2 * R * R, Ri + 2 * R * R + (size_t)I * R};
    const double *cur = U;
    const double *curRi = nullptr;  // Q_0 = U
    for (int k = 0; k < inner; k++) {
        double *nxt = Yb[k & 1];
        DT_TRY(rows(ctx, 0, S, cur, I, R, curRi, nxt, part));
        double *ri = Ri + (size_t)(k & 1) * R * R;
        DT_TRY(launch_chol(ctx, part, nb, R, I, 1, ri, flag));
        cur = nxt;
        curRi = ri;
    }
    // Q = Y R^{-1} (orthonormal to ~cond(Y) u), then one plain Cholesky-QR pass
    DT_TRY(rows(ctx, 1, nullptr, cur, I, R, curRi, U, part));
    double *ri = Ri + (size_t)(inner & 1) * R * R;
    DT_TRY(launch_chol(ctx, part, nb, R, I, 0, ri, flag));
    return rows(ctx, 1, nullptr, U, I, R, ri, U, nullptr, U32, M32, Ub16);
}
