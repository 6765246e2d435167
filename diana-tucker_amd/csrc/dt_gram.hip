// Mode-n Gram S = A_(n) A_(n)^T, cross-Gram A_(n) B_(n)^T and FP64 sums of squares.
//
// Reference: Function::gram(A, n) (template/function/tensor.tpp:18-120) is,
// at one rank, Operator::tenmat (operator_cpu.tpp:56-72) followed by one
// cblas_dgemm NT (operator_double.cpp:176-194); Function::ttt_except
// (tensor.tpp:132-259) is the same with two tensors; Operator<double>::fnorm
// (operator_double.cpp:74-82) is an FP64 reduction.
//
// gram_mfma: symmetric rank-k update on v_mfma_f32_16x16x4_f32.  Workgroup =
// one 64x64 tile of the lower triangle of S and one contiguous range of the
// columns (f, b) of A_(n) (split-K).  Each 32-column chunk of the two row
// panels is staged in LDS (row stride 34 floats: conflict-free MFMA operand
// reads), the four waves each own a 32x32 sub-tile.  Partial tiles of all
// splits are summed in FP64 by gram_reduce, which also mirrors the upper half.
#include "dt_internal.hpp"

#include <mutex>

template <typename T>
__device__ __forceinline__ float ldg_f(const T *p);
template <>
__device__ __forceinline__ float ldg_f<float>(const float *p) { return *p; }
template <>
__device__ __forceinline__ float ldg_f<uint16_t>(const uint16_t *p) { return bf16_to_f32(*p); }

__device__ __forceinline__ void tri_decode(int t, int &ti, int &tj) {
    int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (r * (r + 1) / 2 > t) r--;
    while ((r + 1) * (r + 2) / 2 <= t) r++;
    ti = r;
    tj = t - r * (r + 1) / 2;
}

constexpr int GS = 34;  // LDS row stride (floats)

template <typename TA, bool F1>
__global__ __launch_bounds__(256) void gram_mfma(const TA *__restrict__ A, float *__restrict__ Sp,
                                                 int64_t F, int64_t I, int64_t Ctot,
                                                 int64_t cols_per_split) {
    __shared__ float Xa[64 * GS];
    __shared__ float Xb[64 * GS];
    int ti, tj;
    tri_decode((int)blockIdx.x, ti, tj);
    const bool diag = ti == tj;
    const int64_t c_begin = (int64_t)blockIdx.y * cols_per_split;
    int64_t c_end = c_begin + cols_per_split;
    if (c_end > Ctot) c_end = Ctot;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c = lane & 15;
    const int g = lane >> 4;
    const int wi = (wave >> 1) * 32;
    const int wj = (wave & 1) * 32;

    floatx4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};

    // the chunk starting at c0 into registers: slot q of this thread is row
    // r, column kc of the chunk as in the LDS layout (F1: r = tid % 64,
    // kc = tid / 64 + 4q; else kc = tid % 32, r = tid / 32 + 8q); columns past
    // c_end and rows past I read as zero.  Without F1 the thread's column is
    // the same for every slot, so its (f, b) split is one division per chunk
    // and the slots are a fixed stride of 8 F elements apart.
    const int tid = threadIdx.x;
    const int64_t rowa = (int64_t)ti * 64, rowb = (int64_t)tj * 64;
    auto load = [&](int64_t c0, float (&va)[8], float (&vb)[8]) {
        if (F1) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int64_t kk = c0 + (tid >> 6) + 4 * q;
                const int64_t ia = rowa + (tid & 63), ib = rowb + (tid & 63);
                va[q] = kk < c_end && ia < I ? ldg_f<TA>(A + I * kk + ia) : 0.f;
                vb[q] = !diag && kk < c_end && ib < I ? ldg_f<TA>(A + I * kk + ib) : 0.f;
            }
        } else {
            const int64_t kk = c0 + (tid & 31);
            const bool ok = kk < c_end;
            const int64_t b = !ok ? 0 : (F <= 0x7fffffff && kk <= 0x7fffffff) ? (int64_t)((uint32_t)kk / (uint32_t)F) : kk / F;
            const int64_t base = (kk - b * F) + F * I * b + F * (int64_t)(tid >> 5);
            const TA *pa = A + base + F * rowa, *pb = A + base + F * rowb;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int64_t r = (tid >> 5) + 8 * q;
                va[q] = ok && rowa + r < I ? ldg_f<TA>(pa + 8 * F * q) : 0.f;
                vb[q] = !diag && ok && rowb + r < I ? ldg_f<TA>(pb + 8 * F * q) : 0.f;
            }
        }
    };
    auto stage = [&](const float (&va)[8], const float (&vb)[8]) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int r = F1 ? (tid & 63) : (tid >> 5) + 8 * q, kc = F1 ? (tid >> 6) + 4 * q : (tid & 31);
            Xa[r * GS + kc] = va[q];
            if (!diag) Xb[r * GS + kc] = vb[q];
        }
    };
    auto mma = [&]() {
        const float *XB = diag ? Xa : Xb;
#pragma unroll
        for (int ks = 0; ks < 32; ks += 4) {
            const float a0 = Xa[(wi + c) * GS + ks + g];
            const float a1 = Xa[(wi + 16 + c) * GS + ks + g];
            const float b0 = XB[(wj + c) * GS + ks + g];
            const float b1 = XB[(wj + 16 + c) * GS + ks + g];
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
        }
    };
    // two register sets alternate (the loop body is unrolled by two, so no
    // register holding a load in flight is copied): the loads of chunk k + 1
    // are in flight while chunk k runs from LDS
    float xa[8], xb[8], ya[8], yb[8];
    load(c_begin, xa, xb);
    for (int64_t c0 = c_begin; c0 < c_end; c0 += 64) {
        stage(xa, xb);
        __syncthreads();
        if (c0 + 32 < c_end) load(c0 + 32, ya, yb);
        mma();
        __syncthreads();
        if (c0 + 32 >= c_end) break;
        stage(ya, yb);
        __syncthreads();
        if (c0 + 64 < c_end) load(c0 + 64, xa, xb);
        mma();
        __syncthreads();
    }
    float *S = Sp + (int64_t)blockIdx.y * I * I;
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int nj = 0; nj < 2; nj++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t i = (int64_t)ti * 64 + wi + 16 * mi + 4 * g + v;
                const int64_t j = (int64_t)tj * 64 + wj + 16 * nj + c;
                if (i < I && j < I) S[i + I * j] = acc[mi][nj][v];
            }
}

// gram_big: S = A_(n) A_(n)ᵀ for I >= 256 over many columns.  A workgroup of
// eight waves owns a 256 x 256 block (a, b), a >= b, of 256-row panels of
// A_(n) and one column range (split-K); it stages chunks of GB_KC columns of
// the panel(s) in LDS (two buffers: the next chunk's global loads are in
// flight while the current chunk's MFMAs run, one barrier per chunk) and
// accumulates 16 x 16 tiles on v_mfma_f32_16x16x4_f32.  MFMA step s of lane
// (c, g) pairs k = 4 g + s of the chunk in both operands (the Gram sums over
// k, so any pairing that is the same on both sides gives the same sum), so a
// lane's four values of a 16-row block come from one 16-byte LDS read.
//   diagonal block (a == b): wave w owns the lower tiles of row blocks w and
//     15 - w (17 tiles); off-diagonal: row blocks 2w, 2w + 1 of panel a against
//     the 16 row blocks of panel b (32 tiles).
// Row layout in LDS (pitch GB_KC + 4 floats) when A_(n)'s columns are
// contiguous in memory (F % 32 == 0: a chunk is 32 consecutive f of one b, a
// row is 128 bytes); column layout (pitch rows + 4) for mode 0 (F == 1, a
// column is contiguous).  Rows past I read as zero.
constexpr int GB_T = 512, GB_P = 256, GB_KC = 16, GB_KS = GB_KC + 4, GB_RS = 2 * GB_P + 4;
constexpr int GB_BUF = 2 * GB_P * GB_KS > GB_KC * GB_RS ? 2 * GB_P * GB_KS : GB_KC * GB_RS;

__device__ __forceinline__ floatx4 gb_ld4(const float *p) { return *(const floatx4 *)p; }
__device__ __forceinline__ floatx4 gb_ld4(const uint16_t *p) {
    const uint2 u = *(const uint2 *)p;
    return floatx4{__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u), __uint_as_float(u.y << 16),
                   __uint_as_float(u.y & 0xffff0000u)};
}

template <typename TA, bool KMAJOR>
__global__ __launch_bounds__(GB_T) void gram_big(const TA *__restrict__ A, float *__restrict__ Sp, int64_t F,
                                                 int64_t I, int64_t Ctot, int64_t cps) {
    __shared__ float X[2][GB_BUF];
    int pa, pb;
    tri_decode((int)blockIdx.x, pa, pb);
    const bool diag = pa == pb;
    const int nrows = diag ? GB_P : 2 * GB_P;
    const int64_t c_begin = (int64_t)blockIdx.y * cps;
    const int64_t c_end = c_begin + cps < Ctot ? c_begin + cps : Ctot;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 15, g = lane >> 4;
    // global row of staged row r
    auto grow = [&](int r) -> int64_t { return r < GB_P ? (int64_t)pa * GB_P + r : (int64_t)pb * GB_P + (r - GB_P); };
    // chunk loads: nrows x GB_KC values as float4s, at most 4 per thread
    constexpr int NL = 2 * GB_P * GB_KC / 4 / GB_T;  // 4
    floatx4 ld[NL];
    auto fetch = [&](int64_t c0) {
#pragma unroll
        for (int q = 0; q < NL; q++) {
            const int e = tid + GB_T * q;
            floatx4 v = {0.f, 0.f, 0.f, 0.f};
            if (KMAJOR) {
                // F == 1: element (row, column c0 + k) at A[row + I (c0 + k)]
                const int nr4 = nrows / 4;
                const int k = e / nr4, r4 = e % nr4;
                const int64_t gr = grow(4 * r4);
                if (k < GB_KC && gr < I && c0 + k < c_end) v = gb_ld4(A + gr + I * (c0 + k));
            } else {
                // F % 32 == 0: a chunk is f0 .. f0 + GB_KC - 1 of one b
                const int r = e / (GB_KC / 4), part = e % (GB_KC / 4);
                const int64_t gr = grow(r);
                const int64_t b = c0 / F, f0 = c0 - b * F;
                if (r < nrows && gr < I) v = gb_ld4(A + f0 + 4 * part + F * gr + F * I * b);
            }
            ld[q] = v;
        }
    };
    auto stage = [&](float *buf) {
#pragma unroll
        for (int q = 0; q < NL; q++) {
            const int e = tid + GB_T * q;
            if (KMAJOR) {
                const int nr4 = nrows / 4;
                const int k = e / nr4, r4 = e % nr4;
                if (k < GB_KC) *(floatx4 *)&buf[k * GB_RS + 4 * r4] = ld[q];
            } else {
                const int r = e / (GB_KC / 4), part = e % (GB_KC / 4);
                if (r < nrows) *(floatx4 *)&buf[r * GB_KS + 4 * part] = ld[q];
            }
        }
    };
    // four values of staged row block q for lane (c, g): k = 4 g .. 4 g + 3
    auto frag = [&](const float *buf, int q) -> floatx4 {
        const int r = 16 * q + c;
        if (KMAJOR)
            return floatx4{buf[(4 * g) * GB_RS + r], buf[(4 * g + 1) * GB_RS + r], buf[(4 * g + 2) * GB_RS + r],
                           buf[(4 * g + 3) * GB_RS + r]};
        return *(const floatx4 *)&buf[r * GB_KS + 4 * g];
    };
    floatx4 acc[2][16];
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
        for (int j = 0; j < 16; j++) acc[u][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    // row blocks of this wave's A operands and the first B row block
    const int ra0 = diag ? w : 2 * w, ra1 = diag ? 15 - w : 2 * w + 1, rb0 = diag ? 0 : 16;
    const int lim0 = diag ? w : 15, lim1 = diag ? 15 - w : 15;
    int cur = 0;
    if (c_begin < c_end) {
        fetch(c_begin);
        stage(X[0]);
    }
    __syncthreads();
    for (int64_t c0 = c_begin; c0 < c_end; c0 += GB_KC) {
        const bool more = c0 + GB_KC < c_end;
        if (more) fetch(c0 + GB_KC);
        const float *buf = X[cur];
        const floatx4 a0 = frag(buf, ra0), a1 = frag(buf, ra1);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const floatx4 bj = frag(buf, rb0 + j);
            if (j <= lim0) {
#pragma unroll
                for (int st = 0; st < 4; st++)
                    acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[st], bj[st], acc[0][j], 0, 0, 0);
            }
            if (j <= lim1) {
#pragma unroll
                for (int st = 0; st < 4; st++)
                    acc[1][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[st], bj[st], acc[1][j], 0, 0, 0);
            }
        }
        if (more) stage(X[cur ^ 1]);
        __syncthreads();
        cur ^= 1;
    }
    float *S = Sp + (int64_t)blockIdx.y * I * I;
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int ra = u == 0 ? ra0 : ra1;
        const int lim = u == 0 ? lim0 : lim1;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            if (j > lim) continue;
            const int64_t gj = grow(16 * (rb0 + j) + c);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t gi = grow(16 * ra + 4 * g + v);
                if (gi < I && gj < I) S[gi + I * gj] = acc[u][j][v];
            }
        }
    }
}

// Cross product C = A_(n) B_(n)ᵀ (IA x JB) of two tensors that differ only in
// mode n, on the same MFMA tiling as gram_mfma: workgroup = one 64x64 tile
// (ti over IA, tj over JB) and one column range (split-K).  FP32 partial
// tiles, summed in FP64 by cross_reduce.
template <typename TA, bool F1>
__global__ __launch_bounds__(256) void cross_mfma(const TA *__restrict__ A, const TA *__restrict__ Bt,
                                                  float *__restrict__ Cp, int64_t F, int64_t IA, int64_t JB,
                                                  int64_t Ctot, int64_t cols_per_split, int ntI) {
    __shared__ float Xa[64 * GS];
    __shared__ float Xb[64 * GS];
    const int ti = (int)(blockIdx.x % ntI), tj = (int)(blockIdx.x / ntI);
    const int64_t c_begin = (int64_t)blockIdx.y * cols_per_split;
    int64_t c_end = c_begin + cols_per_split;
    if (c_end > Ctot) c_end = Ctot;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c = lane & 15;
    const int g = lane >> 4;
    const int wi = (wave >> 1) * 32;
    const int wj = (wave & 1) * 32;
    floatx4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int64_t c0 = c_begin; c0 < c_end; c0 += 32) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int e = threadIdx.x + 256 * q;
            int r, kc;
            if (F1) {
                r = e & 63;
                kc = e >> 6;
            } else {
                kc = e & 31;
                r = e >> 5;
            }
            const int64_t kk = c0 + kc;
            float va = 0.f, vb = 0.f;
            if (kk < c_end) {
                int64_t fa, bb;
                if (F1) {
                    fa = 0;
                    bb = kk;
                } else {
                    bb = kk / F;
                    fa = kk - bb * F;
                }
                const int64_t ia = (int64_t)ti * 64 + r;
                const int64_t jb = (int64_t)tj * 64 + r;
                if (ia < IA) va = ldg_f<TA>(A + fa + F * (ia + IA * bb));
                if (jb < JB) vb = ldg_f<TA>(Bt + fa + F * (jb + JB * bb));
            }
            Xa[r * GS + kc] = va;
            Xb[r * GS + kc] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 32; ks += 4) {
            const float a0 = Xa[(wi + c) * GS + ks + g];
            const float a1 = Xa[(wi + 16 + c) * GS + ks + g];
            const float b0 = Xb[(wj + c) * GS + ks + g];
            const float b1 = Xb[(wj + 16 + c) * GS + ks + g];
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    float *Cs = Cp + (int64_t)blockIdx.y * IA * JB;
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int nj = 0; nj < 2; nj++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t i = (int64_t)ti * 64 + wi + 16 * mi + 4 * g + v;
                const int64_t j = (int64_t)tj * 64 + wj + 16 * nj + c;
                if (i < IA && j < JB) Cs[i + IA * j] = acc[mi][nj][v];
            }
}

// The FP64 sum of split q = 0, 1, ... of p[q * total], in that order; the
// partials are loaded eight at a time so that their loads are in flight
// together.
__device__ __forceinline__ double sum_splits(const float *__restrict__ p, int64_t total, int nsplit) {
    double s = 0.0;
    int q = 0;
    for (; q + 8 <= nsplit; q += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = p[(int64_t)(q + u) * total];
#pragma unroll
        for (int u = 0; u < 8; u++) s += (double)v[u];
    }
    for (; q < nsplit; q++) s += (double)p[(int64_t)q * total];
    return s;
}

__global__ void cross_reduce(const float *__restrict__ Cp, double *__restrict__ C, int64_t total, int nsplit) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride)
        C[idx] = sum_splits(Cp + idx, total, nsplit);
}

// S[i, j] = sum over splits of the partial stored in the lower triangle of
// tiles of `tile` rows (64 for gram_mfma, 16 for gram_big).
__global__ void gram_reduce(const float *__restrict__ Sp, double *__restrict__ S, int64_t I,
                            int nsplit, int tile) {
    const int64_t total = I * I;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const bool small = total <= 0x7fffffff;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        int64_t i, j;
        bool lower;
        if (small) {
            const uint32_t u = (uint32_t)idx, n = (uint32_t)I, t = (uint32_t)tile;
            i = u % n;
            j = u / n;
            lower = (uint32_t)i / t >= (uint32_t)j / t;
        } else {
            i = idx % I;
            j = idx / I;
            lower = i / tile >= j / tile;
        }
        const int64_t src = lower ? idx : (j + I * i);
        S[idx] = sum_splits(Sp + src, total, nsplit);
    }
}

// FP64 (and generic) Gram / cross-Gram: one thread per output entry.
//   C[i + IA*j] = sum_{f,b} A[f + F*(i + IA*b)] * Bt[f + F*(j + JB*b)]
template <typename T>
__global__ void cross_gram_generic(const T *__restrict__ A, const T *__restrict__ Bt,
                                   double *__restrict__ C, int64_t F, int64_t IA, int64_t JB,
                                   int64_t Bk) {
    const int64_t total = IA * JB;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const int64_t i = idx % IA, j = idx / IA;
        double s = 0.0;
        for (int64_t b = 0; b < Bk; b++) {
            const T *pa = A + F * (i + IA * b);
            const T *pb = Bt + F * (j + JB * b);
            for (int64_t f = 0; f < F; f++) s += (double)pa[f] * (double)pb[f];
        }
        C[idx] = s;
    }
}

// ---- sums of squares ----------------------------------------------------------
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[16];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; i++) s += red[i];
    }
    __syncthreads();
    return s;
}

template <typename T>
__device__ __forceinline__ double ld_d(const T *p, int64_t i);
template <>
__device__ __forceinline__ double ld_d<double>(const double *p, int64_t i) { return p[i]; }
template <>
__device__ __forceinline__ double ld_d<float>(const float *p, int64_t i) { return (double)p[i]; }
template <>
__device__ __forceinline__ double ld_d<uint16_t>(const uint16_t *p, int64_t i) {
    return (double)bf16_to_f32(p[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void sumsq_kernel(const T *__restrict__ A, int64_t n,
                                                    double *__restrict__ part) {
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = ld_d<T>(A, i);
        s += v * v;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// float4-vectorised variant for FP32 data (n % 4 == 0, 16-byte aligned).
__global__ __launch_bounds__(256) void sumsq_f32x4(const float4 *__restrict__ A, int64_t n4,
                                                   double *__restrict__ part) {
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        const float4 v = A[i];
        s += (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z + (double)v.w * v.w;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// 16-byte-vectorised variant for BF16 data (n % 8 == 0, 16-byte aligned):
// eight values per load, squares exact in FP64.
__global__ __launch_bounds__(256) void sumsq_bf16x8(const uint4 *__restrict__ A, int64_t n8,
                                                    double *__restrict__ part) {
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n8; i += stride) {
        const uint4 v = A[i];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double lo = (double)__uint_as_float(w[j] << 16);
            const double hi = (double)__uint_as_float(w[j] & 0xffff0000u);
            t += lo * lo + hi * hi;
        }
        s += t;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void sum_partials(const double *__restrict__ part, int n, double *__restrict__ out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    s = block_sum(s);
    if (threadIdx.x == 0) out[0] = s;
}

static inline dim3 gs_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 262144) g = 262144;
    if (g < 1) g = 1;
    return dim3((unsigned)g);
}

// Sum of squares, result left on the device in out_dev[0]; block partials in
// workspace slot `slot`.
constexpr int SUMSQ_BLOCKS = 2048;
static int sumsq_to_dev(dt_ctx *ctx, const void *A, int dt, size_t n, double *out_dev, int slot = WS_SUMSQ) {
    const int nb = SUMSQ_BLOCKS;
    double *part = (double *)dt_ws(ctx, slot, nb * sizeof(double));
    if (!part) return DT_ENOMEM;
    if (dt == DT_F32 && (n % 4) == 0 && (((uintptr_t)A) & 15) == 0) {
        DT_LAUNCH("sumsq", sumsq_f32x4, dim3(nb), dim3(256), 0, (const float4 *)A, (int64_t)(n / 4),
                  part);
    } else if (dt == DT_F32) {
        DT_LAUNCH("sumsq", sumsq_kernel<float>, dim3(nb), dim3(256), 0, (const float *)A, (int64_t)n,
                  part);
    } else if (dt == DT_F64) {
        DT_LAUNCH("sumsq", sumsq_kernel<double>, dim3(nb), dim3(256), 0, (const double *)A,
                  (int64_t)n, part);
    } else if (dt == DT_BF16 && (n % 8) == 0 && (((uintptr_t)A) & 15) == 0) {
        DT_LAUNCH("sumsq", sumsq_bf16x8, dim3(nb), dim3(256), 0, (const uint4 *)A, (int64_t)(n / 8), part);
    } else if (dt == DT_BF16) {
        DT_LAUNCH("sumsq", sumsq_kernel<uint16_t>, dim3(nb), dim3(256), 0, (const uint16_t *)A,
                  (int64_t)n, part);
    } else {
        return dt_fail(ctx, DT_EUNSUPPORTED, "sumsq: unsupported type %d", dt);
    }
    DT_LAUNCH("sumsq", sum_partials, dim3(1), dim3(256), 0, part, nb, out_dev);
    return DT_OK;
}

__global__ __launch_bounds__(256) void dot_f64_kernel(const double *__restrict__ a,
                                                      const double *__restrict__ b, int64_t n,
                                                      double *__restrict__ part) {
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        s = fma(a[i], b[i], s);
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

int dt_dot_f64_dev(dt_ctx *ctx, const double *a, const double *b, size_t n, double *out_dev) {
    const int nb = 256;
    double *part = (double *)dt_ws(ctx, WS_SUMSQ, 2048 * sizeof(double));
    if (!part) return DT_ENOMEM;
    DT_LAUNCH("dot", dot_f64_kernel, dim3(nb), dim3(256), 0, a, b, (int64_t)n, part);
    DT_LAUNCH("dot", sum_partials, dim3(1), dim3(256), 0, part, nb, out_dev);
    return DT_OK;
}

int dt_sumsq_dev(dt_ctx *ctx, const void *A, int dt, size_t n, double *out_host) {
    if (n == 0) {
        *out_host = 0.0;
        return DT_OK;
    }
    double *dres = (double *)dt_ws(ctx, WS_STATUS, 64);
    if (!dres) return DT_ENOMEM;
    DT_TRY(sumsq_to_dev(ctx, A, dt, n, dres));
    DT_HIP(hipMemcpyAsync(ctx->pinned, dres, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    *out_host = ctx->pinned[0];
    return DT_OK;
}

int dt_side_sumsq_setup(dt_ctx *ctx, int count, const void *const *src, const int *dt, const size_t *len) {
    DT_TRY(dt_side_join(ctx, nullptr));  // a job left queued is finished first
    DT_CHECK_ARG(count >= 1 && count <= 2, "side work: %d sums", count);
    // buffers now, so that no allocation (and its implicit synchronisation) happens later
    auto &j = ctx->side;
    j.out_dev = (double *)dt_ws(ctx, WS_SIDE_OUT, 64);
    for (int i = 0; i < count; i++)
        if (!dt_ws(ctx, WS_SIDE0 + i, SUMSQ_BLOCKS * sizeof(double))) return DT_ENOMEM;
    if (!j.out_dev) return DT_ENOMEM;
    for (int i = 0; i < count; i++) {
        j.src[i] = src[i];
        j.dt[i] = dt[i];
        j.len[i] = len[i];
    }
    j.count = count;
    j.kicked = false;
    return DT_OK;
}

int dt_side_kick(dt_ctx *ctx) {
    auto &j = ctx->side;
    if (j.count == 0 || j.kicked) return DT_OK;
    j.kicked = true;
    hipStream_t main_stream = ctx->stream;
    const bool timing = ctx->timing;
    if (ctx->side_stream) {
        // the side stream starts after all the main stream holds so far
        DT_HIP(hipEventRecord(ctx->side_fork, main_stream));
        DT_HIP(hipStreamWaitEvent(ctx->side_stream, ctx->side_fork, 0));
        ctx->stream = ctx->side_stream;  // the launches below go to the side stream, untimed
        ctx->timing = false;
    }
    int st = DT_OK;
    for (int i = 0; i < j.count && st == DT_OK; i++)
        st = j.len[i] == 0 ? DT_OK : sumsq_to_dev(ctx, j.src[i], j.dt[i], j.len[i], j.out_dev + i, WS_SIDE0 + i);
    ctx->stream = main_stream;
    ctx->timing = timing;
    if (st == DT_OK && ctx->side_stream) DT_HIP(hipEventRecord(ctx->side_done, ctx->side_stream));
    return st;
}

int dt_side_join(dt_ctx *ctx, double *out_host) {
    auto &j = ctx->side;
    if (j.count == 0) return DT_OK;
    int st = dt_side_kick(ctx);
    if (ctx->side_stream && j.kicked) {
        // after a failed kick side_done may not mark this job's launches: wait
        // for the side stream itself, so no side launch outlives the job
        if (st != DT_OK || hipStreamWaitEvent(ctx->stream, ctx->side_done, 0) != hipSuccess)
            hipStreamSynchronize(ctx->side_stream);
    }
    const int count = j.count;
    j.count = 0;
    j.kicked = false;
    DT_TRY(st);
    if (!out_host) return DT_OK;
    for (int i = 0; i < count; i++)
        if (j.len[i] == 0) DT_HIP(hipMemsetAsync(j.out_dev + i, 0, sizeof(double), ctx->stream));
    DT_HIP(hipMemcpyAsync(ctx->pinned, j.out_dev, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < count; i++) out_host[i] = ctx->pinned[i];
    return DT_OK;
}

extern "C" int dt_sumsq(dt_ctx *ctx, const void *A, int dt, size_t n, double *out_host) {
    DT_CHECK_ARG(ctx && out_host && (A || n == 0), "dt_sumsq: null argument");
    return dt_sumsq_dev(ctx, A, dt, n, out_host);
}

// gram_big applies to orders of at least 256 with many columns, when A_(n)'s
// chunks load as aligned 16-byte (FP32) or 8-byte (bf16) vectors: mode 0 with
// I % 4 == 0, or F % 32 == 0.
static bool gram_big_fits(int64_t F, int64_t I, int64_t Ctot, const void *A) {
    if (I < GB_P || Ctot < (int64_t)1 << 16 || ((uintptr_t)A & 15) != 0) return false;
    return (F == 1 && I % 4 == 0) || (F % 32 == 0);
}

// Split-K plan of the MFMA Gram of order I over Ctot columns, at most
// max_split splits.  gram_mfma: ~512 workgroups of 64 x 64 tiles, at least 256
// columns per split.  gram_big: ~1024 workgroups of 256 x 256 blocks (four per
// CU over the run), at least 1024 columns per split.
static void gram_plan(int64_t I, int64_t Ctot, int64_t max_split, bool big, int64_t &nsplit, int64_t &cps) {
    const int nt = (int)ceil_div(I, big ? GB_P : 64);
    const int ntiles = nt * (nt + 1) / 2;
    nsplit = (big ? 1024 : 512) / ntiles;
    if (nsplit < 1) nsplit = 1;
    int64_t maxsplit = ceil_div(Ctot, big ? 1024 : 256);
    if (maxsplit < 1) maxsplit = 1;
    if (nsplit > maxsplit) nsplit = maxsplit;
    if (nsplit > max_split) nsplit = max_split;
    cps = ceil_div(ceil_div(Ctot, nsplit), 32) * 32;
    nsplit = ceil_div(Ctot, cps);
}

// S (I x I, FP64) from an FP32 / bf16 tensor viewed as [F, I, B] with the
// planned split; Sp holds nsplit I x I FP32 partials.  Every partial entry the
// reduction reads is written by this launch (whole lower-triangle tiles), so
// Sp needs no clearing.
static int gram_mfma_run(dt_ctx *ctx, const void *A, int dtA, int64_t F, int64_t I, int64_t B, double *S,
                         float *Sp, int64_t nsplit, int64_t cps, bool big) {
    const int64_t Ctot = F * B;
    if (big) {
        const int np = (int)ceil_div(I, GB_P);
        dim3 grid((unsigned)(np * (np + 1) / 2), (unsigned)nsplit);
#define GBIG(T, KM) \
    DT_LAUNCH("gram_big", (gram_big<T, KM>), grid, dim3(GB_T), 0, (const T *)A, Sp, F, I, Ctot, cps)
        if (dtA == DT_F32) {
            if (F == 1) {
                GBIG(float, true);
            } else {
                GBIG(float, false);
            }
        } else {
            if (F == 1) {
                GBIG(uint16_t, true);
            } else {
                GBIG(uint16_t, false);
            }
        }
#undef GBIG
        DT_LAUNCH("gram_reduce", gram_reduce, gs_grid(I * I), dim3(256), 0, Sp, S, I, (int)nsplit, 16);
        return DT_OK;
    }
    const int nt = (int)ceil_div(I, 64);
    const int ntiles = nt * (nt + 1) / 2;
    dim3 grid((unsigned)ntiles, (unsigned)nsplit);
    if (dtA == DT_F32) {
        if (F == 1) {
            DT_LAUNCH("gram_mfma", (gram_mfma<float, true>), grid, dim3(256), 0, (const float *)A, Sp,
                      F, I, Ctot, cps);
        } else {
            DT_LAUNCH("gram_mfma", (gram_mfma<float, false>), grid, dim3(256), 0, (const float *)A,
                      Sp, F, I, Ctot, cps);
        }
    } else {
        if (F == 1) {
            DT_LAUNCH("gram_mfma", (gram_mfma<uint16_t, true>), grid, dim3(256), 0,
                      (const uint16_t *)A, Sp, F, I, Ctot, cps);
        } else {
            DT_LAUNCH("gram_mfma", (gram_mfma<uint16_t, false>), grid, dim3(256), 0,
                      (const uint16_t *)A, Sp, F, I, Ctot, cps);
        }
    }
    DT_LAUNCH("gram_reduce", gram_reduce, gs_grid(I * I), dim3(256), 0, Sp, S, I, (int)nsplit, 64);
    return DT_OK;
}

// S (I x I, FP64, device) = A_(n) A_(n)^T for A viewed as [F, I, B].
int dt_gram_internal(dt_ctx *ctx, const void *A, int dtA, int64_t F, int64_t I, int64_t B,
                     double *S) {
    if (I == 0) return DT_OK;
    if (dtA == DT_F64) {
        DT_LAUNCH("gram_generic_f64", cross_gram_generic<double>, gs_grid(I * I), dim3(256), 0,
                  (const double *)A, (const double *)A, S, F, I, I, B);
        return DT_OK;
    }
    DT_CHECK_ARG(dtA == DT_F32 || dtA == DT_BF16, "gram: unsupported type %d", dtA);
    int64_t nsplit, cps;
    const bool big = gram_big_fits(F, I, F * B, A);
    gram_plan(I, F * B, 256, big, nsplit, cps);
    float *Sp = (float *)dt_ws(ctx, WS_GRAM_PART, (size_t)(nsplit * I * I) * sizeof(float));
    if (!Sp) return DT_ENOMEM;
    return gram_mfma_run(ctx, A, dtA, F, I, B, S, Sp, nsplit, cps, big);
}

static int gram_args(dt_ctx *ctx, const void *A, const void *S, const size_t *shape, int ndim, int n, int dtS,
                     const char *who) {
    DT_CHECK_ARG(ctx && A && S && shape, "%s: null argument", who);
    DT_CHECK_ARG(ndim >= 1 && ndim <= DT_MAX_NDIM && n >= 0 && n < ndim, "%s: bad mode", who);
    DT_CHECK_ARG(dtS == DT_F64 || dtS == DT_F32, "%s: output must be F64 or F32", who);
    return DT_OK;
}

extern "C" int dt_gram(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape, int n,
                       void *S, int dtS) {
    DT_TRY(gram_args(ctx, A, S, shape, ndim, n, dtS, "dt_gram"));
    int64_t F = 1, B = 1;
    for (int i = 0; i < n; i++) F *= (int64_t)shape[i];
    for (int i = n + 1; i < ndim; i++) B *= (int64_t)shape[i];
    const int64_t I = (int64_t)shape[n];
    if (dtS == DT_F64) return dt_gram_internal(ctx, A, dtA, F, I, B, (double *)S);
    double *tmp = (double *)dt_ws(ctx, WS_SMALL3, (size_t)(I * I) * sizeof(double));
    if (!tmp) return DT_ENOMEM;
    DT_TRY(dt_gram_internal(ctx, A, dtA, F, I, B, tmp));
    return dt_convert(ctx, S, DT_F32, tmp, DT_F64, (size_t)(I * I));
}

// ---- the Gram without allocations (dt_gram_noalloc) --------------------------
// Its device scratch is this file-scope array: each GPU holds its own copy
// from module load to process exit.  Host code never passes `scratch` itself
// to a kernel (on the host that name is a stand-in, not a device address):
// dt_noalloc_setup, run by dt_create with the context's device current, looks
// up the device address once per device and every later use goes through it.
static __device__ __align__(256) unsigned char scratch[DT_GRAM_NOALLOC_SCRATCH];

constexpr int kNoAllocMaxDevices = 64;
static void *g_scratch_dev[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

int dt_noalloc_setup(dt_ctx *ctx) {
    ctx->noalloc_scratch = nullptr;
    if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices)
        return dt_fail(ctx, DT_EUNSUPPORTED, "allocation-free scratch: device %d >= %d", ctx->device,
                       kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[ctx->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return dt_fail(ctx, DT_EHIP, "allocation-free scratch: hipGetSymbolAddress: %s", hipGetErrorString(e));
        }
        g_scratch_dev[ctx->device] = p;
    }
    ctx->noalloc_scratch = g_scratch_dev[ctx->device];
    return DT_OK;
}

extern "C" int dt_gram_noalloc(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape, int n,
                               void *S, int dtS) {
    DT_TRY(gram_args(ctx, A, S, shape, ndim, n, dtS, "dt_gram_noalloc"));
    DT_CHECK_ARG(dtA == DT_F64 || dtA == DT_F32 || dtA == DT_BF16, "dt_gram_noalloc: type %d", dtA);
    if (!ctx->noalloc_scratch)
        return dt_fail(ctx, DT_EHIP, "dt_gram_noalloc: no scratch address for device %d (%s)", ctx->device,
                       ctx->noalloc_err.c_str());
    int64_t F = 1, B = 1;
    for (int i = 0; i < n; i++) F *= (int64_t)shape[i];
    for (int i = n + 1; i < ndim; i++) B *= (int64_t)shape[i];
    const int64_t I = (int64_t)shape[n];
    if (I == 0) return DT_OK;
    const size_t cap = DT_GRAM_NOALLOC_SCRATCH;
    // an F32 output goes through an FP64 Gram at the start of the scratch
    const size_t tmp_bytes = dtS == DT_F32 ? (((size_t)(I * I) * sizeof(double) + 255) & ~(size_t)255) : 0;
    const size_t part_bytes = dtA == DT_F64 ? 0 : (size_t)(I * I) * sizeof(float);
    if (tmp_bytes + part_bytes > cap)
        return dt_fail(ctx, DT_EINVAL,
                       "dt_gram_noalloc: a Gram of order %lld needs %zu bytes of scratch, above the %zu-byte maximum",
                       (long long)I, tmp_bytes + part_bytes, cap);
    // no event creation inside the entry point: its launches run untimed
    const bool timing = ctx->timing;
    ctx->timing = false;
    unsigned char *base = (unsigned char *)ctx->noalloc_scratch;
    double *Sd = dtS == DT_F64 ? (double *)S : (double *)base;
    int st = DT_OK;
    if (dtA == DT_F64) {
        hipLaunchKernelGGL(cross_gram_generic<double>, gs_grid(I * I), dim3(256), 0, ctx->stream,
                           (const double *)A, (const double *)A, Sd, F, I, I, B);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) st = dt_fail(ctx, DT_EHIP, "dt_gram_noalloc: %s", hipGetErrorString(e));
    } else {
        int64_t nsplit, cps;
        const bool big = gram_big_fits(F, I, F * B, A);
        gram_plan(I, F * B, (int64_t)((cap - tmp_bytes) / part_bytes), big, nsplit, cps);
        st = gram_mfma_run(ctx, A, dtA, F, I, B, Sd, (float *)(base + tmp_bytes), nsplit, cps, big);
    }
    if (st == DT_OK && dtS == DT_F32) st = dt_convert(ctx, S, DT_F32, Sd, DT_F64, (size_t)(I * I));
    ctx->timing = timing;
    return st;
}

// C (IA x JB, FP64) = A_(n) B_(n)ᵀ for FP32 / bf16 tensors (split-K MFMA).
int dt_cross_internal(dt_ctx *ctx, const void *A, const void *B, int dt, int64_t F, int64_t IA, int64_t JB,
                      int64_t Bk, double *C) {
    const int64_t Ctot = F * Bk;
    const int ntI = (int)ceil_div(IA, 64), ntJ = (int)ceil_div(JB, 64);
    const int ntiles = ntI * ntJ;
    int64_t nsplit = 512 / ntiles;
    if (nsplit < 1) nsplit = 1;
    int64_t maxsplit = ceil_div(Ctot, 256);
    if (maxsplit < 1) maxsplit = 1;
    if (nsplit > maxsplit) nsplit = maxsplit;
    if (nsplit > 256) nsplit = 256;
    int64_t cps = ceil_div(ceil_div(Ctot, nsplit), 32) * 32;
    nsplit = ceil_div(Ctot, cps);
    float *Cp = (float *)dt_ws(ctx, WS_TTT, (size_t)(nsplit * IA * JB) * sizeof(float));
    if (!Cp) return DT_ENOMEM;
    dim3 grid((unsigned)ntiles, (unsigned)nsplit);
#define CROSS(T, F1) \
    DT_LAUNCH("cross_mfma", (cross_mfma<T, F1>), grid, dim3(256), 0, (const T *)A, (const T *)B, Cp, F, IA, JB, Ctot, cps, ntI)
    if (dt == DT_F32) {
        if (F == 1) {
            CROSS(float, true);
        } else {
            CROSS(float, false);
        }
    } else {
        if (F == 1) {
            CROSS(uint16_t, true);
        } else {
            CROSS(uint16_t, false);
        }
    }
#undef CROSS
    DT_LAUNCH("cross_reduce", cross_reduce, gs_grid(IA * JB), dim3(256), 0, Cp, C, IA * JB, (int)nsplit);
    return DT_OK;
}

extern "C" int dt_ttt_except(dt_ctx *ctx, const void *A, const void *B, int dt, int ndim,
                             const size_t *shapeA, size_t JB, int n, double *C) {
    DT_CHECK_ARG(ctx && A && B && C && shapeA, "dt_ttt_except: null argument");
    DT_CHECK_ARG(ndim >= 1 && ndim <= DT_MAX_NDIM && n >= 0 && n < ndim, "dt_ttt_except: bad mode");
    int64_t F = 1, Bk = 1;
    for (int i = 0; i < n; i++) F *= (int64_t)shapeA[i];
    for (int i = n + 1; i < ndim; i++) Bk *= (int64_t)shapeA[i];
    const int64_t IA = (int64_t)shapeA[n];
    if (dt == DT_F64) {
        DT_LAUNCH("ttt_except", cross_gram_generic<double>, gs_grid(IA * (int64_t)JB), dim3(256), 0,
                  (const double *)A, (const double *)B, C, F, IA, (int64_t)JB, Bk);
        return DT_OK;
    }
    DT_CHECK_ARG(dt == DT_F32 || dt == DT_BF16, "dt_ttt_except: type %d", dt);
    return dt_cross_internal(ctx, A, B, dt, F, IA, (int64_t)JB, Bk, C);
}
