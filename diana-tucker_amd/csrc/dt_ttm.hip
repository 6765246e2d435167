// Mode-n tensor-times-matrix without unfolding copies.
//
// Reference: Function::ttm (template/function/tensor.tpp:272-335) builds
// A_(n)^T with Operator::tenmatt (operator_cpu.tpp:83-103), multiplies by M^T
// with cblas_dgemm NT (operator_double.cpp:176-194) and folds the result back
// with mattten (operator_cpu.tpp:115-135).  Here the column-major tensor is
// viewed as A[f, k, b] with f < F = prod_{i<n} I_i, k < K = I_n,
// b < B = prod_{i>n} I_i, and the product is taken in place by addressing:
//     Y[f, j, b] = sum_k A[f, k, b] * M[j, k]
// which is exactly tenmatt -> matmulNT -> mattten of the reference.
//
// Kernels (gfx950, wave64, 256 threads = 4 waves per workgroup):
//   ttm_m0_f32 / ttm_m0_bf16  F == 1: Y(J x B) = M (J x K) . A(K x B).  Each lane
//       streams 16-byte k-contiguous slices of A columns straight into the MFMA
//       B operand (v_mfma_f32_16x16x4_f32 / v_mfma_f32_16x16x32_bf16).
//   ttm_mk_f32 / ttm_mk_bf16  F > 1.  Each lane loads 16-byte f-contiguous
//       slices; MFMA number t consumes slice element t, so a 16x16 output block
//       holds f values strided by 4 (fp32) or 8 (bf16) and the stores are
//       16-byte f-contiguous again.  The bf16 variant transposes 8x8 bf16 blocks
//       in registers with v_perm_b32.
//   ttm_generic  any shape and type (F64 path, odd sizes, J > 64).
// All variants accumulate in FP32 (FP64 for the generic FP64 instantiation).
#include "dt_internal.hpp"

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma_f32_4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mfma_bf16_32(u32x4 a, u32x4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

template <typename T>
__device__ __forceinline__ float ld_f(const T *p);
template <>
__device__ __forceinline__ float ld_f<float>(const float *p) { return *p; }
template <>
__device__ __forceinline__ float ld_f<uint16_t>(const uint16_t *p) { return bf16_to_f32(*p); }

// ---------------------------------------------------------------------------
// Generic: one thread per output element.
template <typename TA, typename TM, typename TY, typename TACC>
__global__ __launch_bounds__(256) void ttm_generic(const TA *__restrict__ A, const TM *__restrict__ M,
                                                   TY *__restrict__ Y, int64_t F, int64_t K,
                                                   int64_t B, int64_t J) {
    const int64_t total = F * J * B;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        int64_t f = idx % F;
        int64_t t = idx / F;
        int64_t j = t % J;
        int64_t b = t / J;
        const TA *a = A + f + F * K * b;
        TACC acc = 0;
        for (int64_t k = 0; k < K; k++) {
            TACC av;
            if constexpr (sizeof(TA) == 2)
                av = (TACC)bf16_to_f32(((const uint16_t *)a)[F * k]);
            else
                av = (TACC)a[F * k];
            acc += av * (TACC)M[j + J * k];
        }
        Y[idx] = (TY)acc;
    }
}

// ---------------------------------------------------------------------------
// F == 1, FP32.  Wave = 64 columns of A (4 blocks of 16), all J rows.
//   U: K x J column-major (k fastest).  Requires K % 4 == 0.
template <typename T>
__device__ __forceinline__ T ld_a(const T *p, bool nt) {
    return nt ? __builtin_nontemporal_load(p) : *p;
}

// Every MFMA kernel has a fast path for tiles that are fully inside the tensor
// and a K that is a multiple of the pipelined step: its loads carry no bounds
// checks and run PD steps ahead of the MFMAs through a ring of registers, so
// each wave keeps 16 KB of the streamed tensor in flight.  Edge tiles take the
// guarded loop.

template <int JB, bool NT>
__global__ __launch_bounds__(256) void ttm_m0_f32(const float *__restrict__ A,
                                                  const float *__restrict__ U,
                                                  float *__restrict__ Y, int K, int64_t B, int J) {
    constexpr int PD = 2;  // pipeline depth in 32-k steps
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c = lane & 15;
    const int g = lane >> 4;
    const int64_t bw = ((int64_t)blockIdx.x * 4 + wave) * 64;
    if (bw >= B) return;

    floatx4 acc[JB][4];
#pragma unroll
    for (int jb = 0; jb < JB; jb++)
#pragma unroll
        for (int cb = 0; cb < 4; cb++) acc[jb][cb] = floatx4{0.f, 0.f, 0.f, 0.f};

    // factor operand: clamp the row so the load is unconditional, mask the value
    int jrow[JB];
    float jmask[JB];
#pragma unroll
    for (int jb = 0; jb < JB; jb++) {
        const int j = 16 * jb + c;
        jrow[jb] = j < J ? j : J - 1;
        jmask[jb] = j < J ? 1.f : 0.f;
    }

    const bool fast = (bw + 64 <= B) && (K % (32 * PD) == 0);
    if (fast) {
        const float *Ab = A + (int64_t)K * (bw + c) + 4 * g;
        floatx4 a[PD][4][2];
        floatx4 u[PD][JB][2];
        auto load = [&](int p, int k0) {
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int h = 0; h < 2; h++)
                    a[p][cb][h] = ld_a((const floatx4 *)(Ab + (int64_t)K * 16 * cb + k0 + 16 * h), NT);
#pragma unroll
            for (int jb = 0; jb < JB; jb++)
#pragma unroll
                for (int h = 0; h < 2; h++)
                    u[p][jb][h] = *(const floatx4 *)(U + k0 + 16 * h + 4 * g + (int64_t)K * jrow[jb]) *
                                  jmask[jb];
        };
#pragma unroll
        for (int p = 0; p < PD; p++) load(p, 32 * p);
        auto step = [&](int p) {
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int jb = 0; jb < JB; jb++)
#pragma unroll
                    for (int cb = 0; cb < 4; cb++) {
                        acc[jb][cb] = mfma_f32_4(u[p][jb][h][0], a[p][cb][h][0], acc[jb][cb]);
                        acc[jb][cb] = mfma_f32_4(u[p][jb][h][1], a[p][cb][h][1], acc[jb][cb]);
                        acc[jb][cb] = mfma_f32_4(u[p][jb][h][2], a[p][cb][h][2], acc[jb][cb]);
                        acc[jb][cb] = mfma_f32_4(u[p][jb][h][3], a[p][cb][h][3], acc[jb][cb]);
                    }
        };
        // steady state refills unconditionally; the last PD steps only compute
        int kb = 0;
        for (; kb + 32 * PD < K; kb += 32 * PD) {
#pragma unroll
            for (int p = 0; p < PD; p++) {
                step(p);
                load(p, kb + 32 * (p + PD));
            }
        }
#pragma unroll
        for (int p = 0; p < PD; p++) step(p);
    } else {
        for (int k0 = 0; k0 < K; k0 += 32) {
            floatx4 a[4][2];
#pragma unroll
            for (int cb = 0; cb < 4; cb++) {
                const int64_t b = bw + 16 * cb + c;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int k = k0 + 16 * h + 4 * g;
                    a[cb][h] = (b < B && k < K) ? ld_a((const floatx4 *)(A + k + (int64_t)K * b), NT)
                                                : floatx4{0.f, 0.f, 0.f, 0.f};
                }
            }
            floatx4 u[JB][2];
#pragma unroll
            for (int jb = 0; jb < JB; jb++) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int k = k0 + 16 * h + 4 * g;
                    u[jb][h] = (k < K) ? *(const floatx4 *)(U + k + (int64_t)K * jrow[jb]) * jmask[jb]
                                       : floatx4{0.f, 0.f, 0.f, 0.f};
                }
            }
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int jb = 0; jb < JB; jb++)
#pragma unroll
                    for (int cb = 0; cb < 4; cb++) {
                        acc[jb][cb] = mfma_f32_4(u[jb][h][0], a[cb][h][0], acc[jb][cb]);
                        acc[jb][cb] = mfma_f32_4(u[jb][h][1], a[cb][h][1], acc[jb][cb]);
                        acc[jb][cb] = mfma_f32_4(u[jb][h][2], a[cb][h][2], acc[jb][cb]);
                        acc[jb][cb] = mfma_f32_4(u[jb][h][3], a[cb][h][3], acc[jb][cb]);
                    }
        }
    }
    // D[i][col]: i = 4g + v (row of the 16-block), col = c.
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        const int64_t b = bw + 16 * cb + c;
        if (b >= B) continue;
#pragma unroll
        for (int jb = 0; jb < JB; jb++) {
            const int j = 16 * jb + 4 * g;
            float *y = Y + j + (int64_t)J * b;
            if ((J & 3) == 0 && j + 3 < J) {
                *(float4 *)y = make_float4(acc[jb][cb][0], acc[jb][cb][1], acc[jb][cb][2],
                                           acc[jb][cb][3]);
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++)
                    if (j + v < J) y[v] = acc[jb][cb][v];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// F > 1, FP32.  Wave = FB*64 consecutive f for one b, all J columns.
//   M: J x K column-major (j fastest).  Requires F % 4 == 0.
template <int JB, int FB, bool NT>
__global__ __launch_bounds__(256) void ttm_mk_f32(const float *__restrict__ A,
                                                  const float *__restrict__ M,
                                                  float *__restrict__ Y, int64_t F, int K,
                                                  int64_t B, int J, int64_t tiles_per_b) {
    constexpr int PD = FB >= 4 ? 4 : 8;  // pipeline depth in 4-k steps
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c = lane & 15;
    const int g = lane >> 4;
    const int64_t w = (int64_t)blockIdx.x * 4 + wave;
    const int64_t b = w / tiles_per_b;
    if (b >= B) return;
    const int64_t f0 = (w - b * tiles_per_b) * (64 * FB);
    const float *Ab = A + F * (int64_t)K * b;

    floatx4 acc[JB][FB][4];
#pragma unroll
    for (int jb = 0; jb < JB; jb++)
#pragma unroll
        for (int fb = 0; fb < FB; fb++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[jb][fb][t] = floatx4{0.f, 0.f, 0.f, 0.f};

    int jcol[JB];
    float jmask[JB];
#pragma unroll
    for (int jb = 0; jb < JB; jb++) {
        const int j = 16 * jb + c;
        jcol[jb] = j < J ? j : J - 1;
        jmask[jb] = j < J ? 1.f : 0.f;
    }

    const bool fast = (f0 + 64 * FB <= F) && (K % (4 * PD) == 0);
    if (fast) {
        const float *Ap = Ab + f0 + 4 * c + F * g;
        floatx4 a[PD][FB];
        float u[PD][JB];
        auto load = [&](int p, int k0) {
#pragma unroll
            for (int fb = 0; fb < FB; fb++)
                a[p][fb] = ld_a((const floatx4 *)(Ap + 64 * fb + F * (int64_t)k0), NT);
#pragma unroll
            for (int jb = 0; jb < JB; jb++) u[p][jb] = M[jcol[jb] + (int64_t)J * (k0 + g)] * jmask[jb];
        };
#pragma unroll
        for (int p = 0; p < PD; p++) load(p, 4 * p);
        auto step = [&](int p) {
#pragma unroll
            for (int jb = 0; jb < JB; jb++)
#pragma unroll
                for (int fb = 0; fb < FB; fb++) {
                    acc[jb][fb][0] = mfma_f32_4(u[p][jb], a[p][fb][0], acc[jb][fb][0]);
                    acc[jb][fb][1] = mfma_f32_4(u[p][jb], a[p][fb][1], acc[jb][fb][1]);
                    acc[jb][fb][2] = mfma_f32_4(u[p][jb], a[p][fb][2], acc[jb][fb][2]);
                    acc[jb][fb][3] = mfma_f32_4(u[p][jb], a[p][fb][3], acc[jb][fb][3]);
                }
        };
        int kb = 0;
        for (; kb + 4 * PD < K; kb += 4 * PD) {
#pragma unroll
            for (int p = 0; p < PD; p++) {
                step(p);
                load(p, kb + 4 * (p + PD));
            }
        }
#pragma unroll
        for (int p = 0; p < PD; p++) step(p);
    } else {
        for (int k0 = 0; k0 < K; k0 += 4) {
            const int k = k0 + g;
            floatx4 a[FB];
#pragma unroll
            for (int fb = 0; fb < FB; fb++) {
                const int64_t f = f0 + 64 * fb + 4 * c;
                a[fb] = (k < K && f < F) ? ld_a((const floatx4 *)(Ab + f + F * k), NT)
                                         : floatx4{0.f, 0.f, 0.f, 0.f};
            }
            float u[JB];
#pragma unroll
            for (int jb = 0; jb < JB; jb++) u[jb] = (k < K) ? M[jcol[jb] + (int64_t)J * k] * jmask[jb] : 0.f;
#pragma unroll
            for (int jb = 0; jb < JB; jb++)
#pragma unroll
                for (int fb = 0; fb < FB; fb++) {
                    acc[jb][fb][0] = mfma_f32_4(u[jb], a[fb][0], acc[jb][fb][0]);
                    acc[jb][fb][1] = mfma_f32_4(u[jb], a[fb][1], acc[jb][fb][1]);
                    acc[jb][fb][2] = mfma_f32_4(u[jb], a[fb][2], acc[jb][fb][2]);
                    acc[jb][fb][3] = mfma_f32_4(u[jb], a[fb][3], acc[jb][fb][3]);
                }
        }
    }
    // acc[jb][fb][t][v] = Y[f0 + 64fb + 4c + t, 16jb + 4g + v, b]
#pragma unroll
    for (int fb = 0; fb < FB; fb++) {
        const int64_t f = f0 + 64 * fb + 4 * c;
        if (f >= F) continue;
#pragma unroll
        for (int jb = 0; jb < JB; jb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = 16 * jb + 4 * g + v;
                if (j < J)
                    *(float4 *)(Y + f + F * (j + (int64_t)J * b)) =
                        make_float4(acc[jb][fb][0][v], acc[jb][fb][1][v], acc[jb][fb][2][v],
                                    acc[jb][fb][3][v]);
            }
    }
}

// ---------------------------------------------------------------------------
// F == 1, BF16 input, FP32 accumulate/output.  Ub: K x J bf16 (k fastest).
// Requires K % 8 == 0.  Wave = 64 columns, k-step 64 (two 128-byte halves of
// each column line per lane group).
template <int JB, bool NT>
__global__ __launch_bounds__(256) void ttm_m0_bf16(const uint16_t *__restrict__ A,
                                                   const uint16_t *__restrict__ Ub,
                                                   float *__restrict__ Y, int K, int64_t B, int J) {
    constexpr int PD = 2;  // pipeline depth in 64-k steps
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c = lane & 15;
    const int g = lane >> 4;
    const int64_t bw = ((int64_t)blockIdx.x * 4 + wave) * 64;
    if (bw >= B) return;

    floatx4 acc[JB][4];
#pragma unroll
    for (int jb = 0; jb < JB; jb++)
#pragma unroll
        for (int cb = 0; cb < 4; cb++) acc[jb][cb] = floatx4{0.f, 0.f, 0.f, 0.f};
    int jrow[JB];
    bool jin[JB];
#pragma unroll
    for (int jb = 0; jb < JB; jb++) {
        const int j = 16 * jb + c;
        jrow[jb] = j < J ? j : J - 1;
        jin[jb] = j < J;
    }
    const u32x4 zero4 = {0u, 0u, 0u, 0u};
    const bool fast = (bw + 64 <= B) && (K % (64 * PD) == 0);
    if (fast) {
        const uint16_t *Ab = A + (int64_t)K * (bw + c) + 8 * g;
        u32x4 a[PD][4][2], u[PD][JB][2];
        auto load = [&](int p, int k0) {
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int h = 0; h < 2; h++)
                    a[p][cb][h] = ld_a((const u32x4 *)(Ab + (int64_t)K * 16 * cb + k0 + 32 * h), NT);
#pragma unroll
            for (int jb = 0; jb < JB; jb++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const u32x4 v = *(const u32x4 *)(Ub + k0 + 32 * h + 8 * g + (int64_t)K * jrow[jb]);
                    u[p][jb][h] = jin[jb] ? v : zero4;
                }
        };
#pragma unroll
        for (int p = 0; p < PD; p++) load(p, 64 * p);
        auto step = [&](int p) {
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int jb = 0; jb < JB; jb++)
#pragma unroll
                    for (int cb = 0; cb < 4; cb++)
                        acc[jb][cb] = mfma_bf16_32(u[p][jb][h], a[p][cb][h], acc[jb][cb]);
        };
        int kb = 0;
        for (; kb + 64 * PD < K; kb += 64 * PD) {
#pragma unroll
            for (int p = 0; p < PD; p++) {
                step(p);
                load(p, kb + 64 * (p + PD));
            }
        }
#pragma unroll
        for (int p = 0; p < PD; p++) step(p);
    } else {
        for (int k0 = 0; k0 < K; k0 += 64) {
            u32x4 a[4][2];
#pragma unroll
            for (int cb = 0; cb < 4; cb++) {
                const int64_t b = bw + 16 * cb + c;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int k = k0 + 32 * h + 8 * g;
                    a[cb][h] = (b < B && k < K) ? ld_a((const u32x4 *)(A + k + (int64_t)K * b), NT) : zero4;
                }
            }
            u32x4 u[JB][2];
#pragma unroll
            for (int jb = 0; jb < JB; jb++) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int k = k0 + 32 * h + 8 * g;
                    u[jb][h] = (jin[jb] && k < K) ? *(const u32x4 *)(Ub + k + (int64_t)K * jrow[jb]) : zero4;
                }
            }
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int jb = 0; jb < JB; jb++)
#pragma unroll
                    for (int cb = 0; cb < 4; cb++)
                        acc[jb][cb] = mfma_bf16_32(u[jb][h], a[cb][h], acc[jb][cb]);
        }
    }
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        const int64_t b = bw + 16 * cb + c;
        if (b >= B) continue;
#pragma unroll
        for (int jb = 0; jb < JB; jb++) {
            const int j = 16 * jb + 4 * g;
            float *y = Y + j + (int64_t)J * b;
            if ((J & 3) == 0 && j + 3 < J) {
                *(float4 *)y = make_float4(acc[jb][cb][0], acc[jb][cb][1], acc[jb][cb][2],
                                           acc[jb][cb][3]);
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++)
                    if (j + v < J) y[v] = acc[jb][cb][v];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// F == 1, BF16 input, the factor held in LDS (K (16 JB) bf16 <= 128 KB).  The
// operand layout of v_mfma_f32_16x16x32_bf16 puts 16 columns x 64 bytes in
// one load (half of each 128-byte line, like a random row gather); here each
// load instruction takes 8 columns x 128 bytes (lane L: column L / 8, bytes
// 16 (L % 8)), the wave stores its 32 columns of the step to its own LDS tile
// (column stride 144 bytes, so a 16-lane group of the operand reads below
// covers all 64 banks) and reads them back in the operand layout.  The
// factor, pre-arranged in operand order (m0_factor_swizzle), is loaded into
// LDS once per workgroup.  Persistent: wave w of workgroup b takes column
// groups (32 columns) 4 b + w, + 4 G, ... (G workgroups) as one stream of
// 64-k steps, PD of them in flight, so the vector-memory queue holds only the
// tensor's loads and never drains at a group boundary.
constexpr int M0L_CST = 144;         // LDS bytes per column of a wave's tile
constexpr int M0L_UMAX = 8192;       // u32x4 of the factor in LDS (128 KB): K JB <= 4096

// Um[((s * JB + jb) * 2 + h) * 64 + lane] = U rows k = 64 s + 32 h + 8 g .. + 7
// of column j = 16 jb + c as 8 bf16 (zero for j >= J), lane = 16 g + c.
__global__ void m0_factor_swizzle(const uint16_t *__restrict__ Ub, int K, int J, int JB, u32x4 *__restrict__ Um) {
    const int64_t total = (int64_t)(K / 64) * JB * 2 * 64;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int lane = (int)(e & 63);
        const int64_t q = e >> 6;
        const int h = (int)(q & 1);
        const int jb = (int)((q >> 1) % JB);
        const int s = (int)((q >> 1) / JB);
        const int c = lane & 15, g = lane >> 4;
        const int j = 16 * jb + c;
        u32x4 v = {0u, 0u, 0u, 0u};
        if (j < J) v = *(const u32x4 *)(Ub + (int64_t)K * j + 64 * s + 32 * h + 8 * g);
        Um[e] = v;
    }
}

// B % 32 == 0 (ngroups = B / 32), K % (64 PD) == 0, K JB <= 4096.
template <int JB, int PD, bool NT>
__global__ __launch_bounds__(256) void ttm_m0_bf16_lds(const uint16_t *__restrict__ A,
                                                       const u32x4 *__restrict__ Um,
                                                       float *__restrict__ Y, int K, int64_t ngroups, int J) {
    constexpr int CB = 2;            // 16-column blocks per group
    constexpr int NL = 2 * CB;       // load instructions per step (8 columns each)
    constexpr int USL = JB * 2 * 64; // u32x4 per step of the factor
    __shared__ u32x4 Us[M0L_UMAX];
    __shared__ __attribute__((aligned(16))) unsigned char tile[4][16 * CB * M0L_CST];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 15;
    const int g = lane >> 4;
    const int nsteps = K / 64;
    for (int e = tid; e < nsteps * USL; e += 256) Us[e] = Um[e];
    __syncthreads();
    const int64_t gstride = (int64_t)gridDim.x * 4;
    const int64_t g0 = (int64_t)blockIdx.x * 4 + wave;
    if (g0 >= ngroups) return;  // after the only workgroup barrier
    unsigned char *T = tile[wave];
    const int64_t mygroups = (ngroups - g0 + gstride - 1) / gstride;

    // load cursor (group lg, step ls), PD steps ahead of the compute cursor;
    // past the wave's last group it reloads its first group (not used)
    const int lc = lane >> 3, lq = lane & 7;
    const uint32_t off0 = (uint32_t)(K * lc + 8 * lq);
    int64_t lg = g0;
    int ls = 0;
    u32x4 r[PD][NL];
    auto load = [&](int p) {
        const int64_t gg = lg < ngroups ? lg : g0;
        const uint16_t *Aw = A + (int64_t)K * (32 * gg);
#pragma unroll
        for (int i = 0; i < NL; i++)
            r[p][i] = ld_a((const u32x4 *)(Aw + (off0 + (uint32_t)(8 * i * K + 64 * ls))), NT);
        const bool wrap = ls + 1 == nsteps;  // selects, no branch: the wait counts stay exact
        lg = wrap ? lg + gstride : lg;
        ls = wrap ? 0 : ls + 1;
    };
#pragma unroll
    for (int p = 0; p < PD; p++) load(p);
    floatx4 acc[JB][CB];
#pragma unroll
    for (int jb = 0; jb < JB; jb++)
#pragma unroll
        for (int cb = 0; cb < CB; cb++) acc[jb][cb] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int64_t t = 0; t < mygroups; t++) {
        const int64_t cg = g0 + t * gstride;
        for (int s0 = 0; s0 < nsteps; s0 += PD) {
#pragma unroll
            for (int p = 0; p < PD; p++) {
                const int s = s0 + p;
#pragma unroll
                for (int i = 0; i < NL; i++) *(u32x4 *)(T + (8 * i + lc) * M0L_CST + 16 * lq) = r[p][i];
                load(p);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                u32x4 a[2][CB], u[2][JB];
#pragma unroll
                for (int h = 0; h < 2; h++) {
#pragma unroll
                    for (int jb = 0; jb < JB; jb++) u[h][jb] = Us[(s * JB + jb) * 128 + h * 64 + lane];
#pragma unroll
                    for (int cb = 0; cb < CB; cb++)
                        a[h][cb] = *(const u32x4 *)(T + (16 * cb + c) * M0L_CST + 16 * (4 * h + g));
                }
#pragma unroll
                for (int h = 0; h < 2; h++)
#pragma unroll
                    for (int jb = 0; jb < JB; jb++)
#pragma unroll
                        for (int cb = 0; cb < CB; cb++) acc[jb][cb] = mfma_bf16_32(u[h][jb], a[h][cb], acc[jb][cb]);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        // group done: its 32 columns of Y, then a fresh accumulator
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            const int64_t b = 32 * cg + 16 * cb + c;
#pragma unroll
            for (int jb = 0; jb < JB; jb++) {
                const int j = 16 * jb + 4 * g;
                float *y = Y + j + (int64_t)J * b;
                if ((J & 3) == 0 && j + 3 < J) {
                    *(float4 *)y = make_float4(acc[jb][cb][0], acc[jb][cb][1], acc[jb][cb][2], acc[jb][cb][3]);
                } else {
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        if (j + v < J) y[v] = acc[jb][cb][v];
                }
                acc[jb][cb] = floatx4{0.f, 0.f, 0.f, 0.f};
            }
        }
    }
}

// ---------------------------------------------------------------------------
// F > 1, BF16 input.  Wave = 128 consecutive f (8 per lane column) for one b.
// Each lane loads 8 k-rows of 8 f-contiguous bf16 and transposes the 8x8 block
// in registers so that MFMA t receives the 8 k values of f = f0 + 8c + t.
// Requires F % 8 == 0 and K % 8 == 0.
template <int JB>
__device__ __forceinline__ void mk_bf16_step(const u32x4 (&r)[8], const u32x4 (&u)[JB],
                                             floatx4 (&acc)[JB][8]) {
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const uint32_t sel = (t & 1) ? 0x07060302u : 0x05040100u;
        u32x4 bv;
        bv.x = __builtin_amdgcn_perm(r[1][t >> 1], r[0][t >> 1], sel);
        bv.y = __builtin_amdgcn_perm(r[3][t >> 1], r[2][t >> 1], sel);
        bv.z = __builtin_amdgcn_perm(r[5][t >> 1], r[4][t >> 1], sel);
        bv.w = __builtin_amdgcn_perm(r[7][t >> 1], r[6][t >> 1], sel);
#pragma unroll
        for (int jb = 0; jb < JB; jb++) acc[jb][t] = mfma_bf16_32(u[jb], bv, acc[jb][t]);
    }
}

template <int JB, bool NT>
__global__ __launch_bounds__(256) void ttm_mk_bf16(const uint16_t *__restrict__ A,
                                                   const uint16_t *__restrict__ Ub,
                                                   float *__restrict__ Y, int64_t F, int K,
                                                   int64_t B, int J, int64_t tiles_per_b) {
    constexpr int PD = 2;  // pipeline depth in 32-k steps
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c = lane & 15;
    const int g = lane >> 4;
    const int64_t w = (int64_t)blockIdx.x * 4 + wave;
    const int64_t b = w / tiles_per_b;
    if (b >= B) return;
    const int64_t f = (w - b * tiles_per_b) * 128 + 8 * c;
    const uint16_t *Ab = A + F * (int64_t)K * b;
    const bool fin = f < F;

    floatx4 acc[JB][8];
#pragma unroll
    for (int jb = 0; jb < JB; jb++)
#pragma unroll
        for (int t = 0; t < 8; t++) acc[jb][t] = floatx4{0.f, 0.f, 0.f, 0.f};
    int jrow[JB];
    bool jin[JB];
#pragma unroll
    for (int jb = 0; jb < JB; jb++) {
        const int j = 16 * jb + c;
        jrow[jb] = j < J ? j : J - 1;
        jin[jb] = j < J;
    }
    const u32x4 zero4 = {0u, 0u, 0u, 0u};
    const int64_t ftile0 = (w - b * tiles_per_b) * 128;
    const bool fast = (ftile0 + 128 <= F) && (K % (32 * PD) == 0);
    if (fast) {
        const uint16_t *Ap = Ab + f + F * (int64_t)(8 * g);
        u32x4 r[PD][8], u[PD][JB];
        auto load = [&](int p, int k0) {
#pragma unroll
            for (int e = 0; e < 8; e++) r[p][e] = ld_a((const u32x4 *)(Ap + F * (int64_t)(k0 + e)), NT);
#pragma unroll
            for (int jb = 0; jb < JB; jb++) {
                const u32x4 v = *(const u32x4 *)(Ub + k0 + 8 * g + (int64_t)K * jrow[jb]);
                u[p][jb] = jin[jb] ? v : zero4;
            }
        };
#pragma unroll
        for (int p = 0; p < PD; p++) load(p, 32 * p);
        int kb = 0;
        for (; kb + 32 * PD < K; kb += 32 * PD) {
#pragma unroll
            for (int p = 0; p < PD; p++) {
                mk_bf16_step<JB>(r[p], u[p], acc);
                load(p, kb + 32 * (p + PD));
            }
        }
#pragma unroll
        for (int p = 0; p < PD; p++) mk_bf16_step<JB>(r[p], u[p], acc);
    } else {
        for (int k0 = 0; k0 < K; k0 += 32) {
            const int kb = k0 + 8 * g;
            const bool kin = kb < K;
            u32x4 r[8];
#pragma unroll
            for (int e = 0; e < 8; e++)
                r[e] = (kin && fin) ? ld_a((const u32x4 *)(Ab + f + F * (int64_t)(kb + e)), NT) : zero4;
            u32x4 u[JB];
#pragma unroll
            for (int jb = 0; jb < JB; jb++)
                u[jb] = (kin && jin[jb]) ? *(const u32x4 *)(Ub + kb + (int64_t)K * jrow[jb]) : zero4;
            mk_bf16_step<JB>(r, u, acc);
        }
    }
    if (!fin) return;
#pragma unroll
    for (int jb = 0; jb < JB; jb++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int j = 16 * jb + 4 * g + v;
            if (j >= J) continue;
            float *y = Y + f + F * (j + (int64_t)J * b);
            *(float4 *)y = make_float4(acc[jb][0][v], acc[jb][1][v], acc[jb][2][v], acc[jb][3][v]);
            *(float4 *)(y + 4) =
                make_float4(acc[jb][4][v], acc[jb][5][v], acc[jb][6][v], acc[jb][7][v]);
        }
}

// ---------------------------------------------------------------------------
// FP64-accumulating mode product for the accurate core: one thread per (f, b),
// 16 output columns j per pass, A read as FP32 or BF16 and widened to FP64.
template <typename TA>
__global__ __launch_bounds__(256) void ttm_f64acc(const TA *__restrict__ A,
                                                  const double *__restrict__ M,
                                                  double *__restrict__ Y, int64_t F, int64_t K,
                                                  int64_t B, int64_t J) {
    const int64_t total = F * B;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const int64_t f = idx % F, b = idx / F;
        const TA *a = A + f + F * K * b;
        for (int64_t j0 = 0; j0 < J; j0 += 16) {
            double acc[16];
#pragma unroll
            for (int q = 0; q < 16; q++) acc[q] = 0.0;
            const int jn = (int)((J - j0) < 16 ? (J - j0) : 16);
            for (int64_t k = 0; k < K; k++) {
                double av;
                if constexpr (sizeof(TA) == 2)
                    av = (double)bf16_to_f32(((const uint16_t *)a)[F * k]);
                else
                    av = (double)a[F * k];
                const double *m = M + j0 + J * k;
#pragma unroll
                for (int q = 0; q < 16; q++)
                    if (q < jn) acc[q] = fma(av, m[q], acc[q]);
            }
            for (int q = 0; q < jn; q++) Y[f + F * ((j0 + q) + J * b)] = acc[q];
        }
    }
}

int dt_ttm_f64acc(dt_ctx *ctx, const void *A, int dtA, int64_t F, int64_t K, int64_t B,
                  const double *M64, int64_t J, double *Y) {
    const int64_t total = F * B;
    int64_t g = (total + 255) / 256;
    if (g > 262144) g = 262144;
    if (g < 1) g = 1;
    if (dtA == DT_F32) {
        DT_LAUNCH("ttm_f64acc", ttm_f64acc<float>, dim3((unsigned)g), dim3(256), 0, (const float *)A,
                  M64, Y, F, K, B, J);
    } else if (dtA == DT_BF16) {
        DT_LAUNCH("ttm_f64acc", ttm_f64acc<uint16_t>, dim3((unsigned)g), dim3(256), 0,
                  (const uint16_t *)A, M64, Y, F, K, B, J);
    } else {
        DT_LAUNCH("ttm_f64acc", ttm_f64acc<double>, dim3((unsigned)g), dim3(256), 0, (const double *)A,
                  M64, Y, F, K, B, J);
    }
    return DT_OK;
}

// ---------------------------------------------------------------------------
// Transpose with conversion: dst (n x m) = src (m x n)^T, column-major.
template <typename TD, typename TS>
__global__ void transpose_convert(TD *__restrict__ dst, const TS *__restrict__ src, int64_t m,
                                  int64_t n) {
    const int64_t total = m * n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        int64_t i = idx % m, j = idx / m;  // src(i, j)
        double v = (double)src[idx];
        if constexpr (sizeof(TD) == 2)
            dst[j + n * i] = f32_to_bf16((float)v);
        else
            dst[j + n * i] = (TD)v;
    }
}
template <typename TD, typename TS>
__global__ void copy_convert(TD *__restrict__ dst, const TS *__restrict__ src, int64_t total) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        double v = (double)src[idx];
        if constexpr (sizeof(TD) == 2)
            dst[idx] = f32_to_bf16((float)v);
        else
            dst[idx] = (TD)v;
    }
}

static inline dim3 gs_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 262144) g = 262144;
    if (g < 1) g = 1;
    return dim3((unsigned)g);
}

int dt_transpose_f64(dt_ctx *ctx, double *B, const double *A, size_t m, size_t n) {
    DT_LAUNCH("transpose", (transpose_convert<double, double>), gs_grid(m * n), dim3(256), 0, B, A,
              (int64_t)m, (int64_t)n);
    return DT_OK;
}

// From an FP64 K x J factor build the kernel views (any pointer may be null).
int dt_f64_to_views(dt_ctx *ctx, const double *U64, size_t K, size_t J, float *U32, float *M32,
                    uint16_t *Ub16, double *M64) {
    const int64_t n = (int64_t)(K * J);
    if (U32) {
        DT_LAUNCH("factor_views", (copy_convert<float, double>), gs_grid(n), dim3(256), 0, U32, U64, n);
    }
    if (M32) {
        DT_LAUNCH("factor_views", (transpose_convert<float, double>), gs_grid(n), dim3(256), 0, M32,
                  U64, (int64_t)K, (int64_t)J);
    }
    if (Ub16) {
        DT_LAUNCH("factor_views", (copy_convert<uint16_t, double>), gs_grid(n), dim3(256), 0, Ub16,
                  U64, n);
    }
    if (M64) {
        DT_LAUNCH("factor_views", (transpose_convert<double, double>), gs_grid(n), dim3(256), 0, M64,
                  U64, (int64_t)K, (int64_t)J);
    }
    return DT_OK;
}

// ---------------------------------------------------------------------------
int dt_ttm_views(dt_ctx *ctx, const void *A, int dtA, int64_t F, int64_t K, int64_t B,
                 const dt_factor_views &fv, int64_t J, void *Y, int dtY, bool stream) {
    const int64_t total = F * J * B;
    if (total == 0) return DT_OK;
    if (dtA == DT_F64) {
        DT_CHECK_ARG(dtY == DT_F64 && fv.M64, "ttm: F64 input needs F64 output and factor");
        DT_LAUNCH("ttm_generic_f64", (ttm_generic<double, double, double, double>), gs_grid(total),
                  dim3(256), 0, (const double *)A, fv.M64, (double *)Y, F, K, B, J);
        return DT_OK;
    }
    DT_CHECK_ARG(dtY == DT_F32, "ttm: F32/BF16 input needs F32 output");
    const int jbn = (int)ceil_div(J, 16);
    if (dtA == DT_F32) {
        if (J <= 64 && F == 1 && (K % 4) == 0 && fv.U32) {
            dim3 grid((unsigned)ceil_div(B, 256));
#define M0F(JBV)                                                                          \
    if (stream) {                                                                         \
        DT_LAUNCH("ttm_m0_f32_stream", (ttm_m0_f32<JBV, true>), grid, dim3(256), 0,       \
                  (const float *)A, fv.U32, (float *)Y, (int)K, B, (int)J);               \
    } else {                                                                              \
        DT_LAUNCH("ttm_m0_f32", (ttm_m0_f32<JBV, false>), grid, dim3(256), 0,             \
                  (const float *)A, fv.U32, (float *)Y, (int)K, B, (int)J);               \
    }                                                                                     \
    return DT_OK;
            if (jbn == 1) { M0F(1) }
            if (jbn == 2) { M0F(2) }
            M0F(4)
#undef M0F
        }
        if (J <= 64 && F > 1 && (F % 4) == 0 && fv.M32) {
            // f tile per wave: 64*FB; keep JB*FB <= 4, and take a wider tile
            // only while it still leaves >= 2048 waves (8 per CU) to keep
            // enough loads in flight (a 67 MB tensor at FB = 4 is 256 waves)
            auto waves_at = [&](int fbv) { return ceil_div(F, 64 * fbv) * B; };
            int fb = 1;
            if (jbn == 1) fb = F >= 256 && waves_at(4) >= 2048 ? 4 : (F >= 128 && waves_at(2) >= 2048 ? 2 : 1);
            else if (jbn == 2) fb = F >= 128 && waves_at(2) >= 2048 ? 2 : 1;
            const int64_t tiles = ceil_div(F, 64 * fb);
            const int64_t waves = tiles * B;
            dim3 grid((unsigned)ceil_div(waves, 4));
#define MKF(JBV, FBV)                                                                     \
    if (stream) {                                                                         \
        DT_LAUNCH("ttm_mk_f32_stream", (ttm_mk_f32<JBV, FBV, true>), grid, dim3(256), 0,  \
                  (const float *)A, fv.M32, (float *)Y, F, (int)K, B, (int)J, tiles);     \
    } else {                                                                              \
        DT_LAUNCH("ttm_mk_f32", (ttm_mk_f32<JBV, FBV, false>), grid, dim3(256), 0,        \
                  (const float *)A, fv.M32, (float *)Y, F, (int)K, B, (int)J, tiles);     \
    }                                                                                     \
    return DT_OK;
            if (jbn == 1 && fb == 4) { MKF(1, 4) }
            if (jbn == 1 && fb == 2) { MKF(1, 2) }
            if (jbn == 1) { MKF(1, 1) }
            if (jbn == 2 && fb == 2) { MKF(2, 2) }
            if (jbn == 2) { MKF(2, 1) }
            MKF(4, 1)
#undef MKF
        }
        DT_CHECK_ARG(fv.M32, "ttm: missing FP32 factor");
        DT_LAUNCH("ttm_generic_f32", (ttm_generic<float, float, float, float>), gs_grid(total),
                  dim3(256), 0, (const float *)A, fv.M32, (float *)Y, F, K, B, J);
        return DT_OK;
    }
    if (dtA == DT_BF16) {
        // whole-line loads through LDS with the factor resident there: B % 32 == 0,
        // K % 128 == 0 and the factor's operand copy within 128 KB
        const int jbv = jbn == 1 ? 1 : (jbn == 2 ? 2 : 4);
        if (J <= 64 && F == 1 && (K % 128) == 0 && B % 32 == 0 && (int64_t)K * jbv <= 4096 && fv.Ub16) {
            u32x4 *Um = (u32x4 *)dt_ws(ctx, WS_TTM_SWZ, (size_t)K * 16 * jbv * sizeof(uint16_t));
            if (!Um) return DT_ENOMEM;
            const int64_t nsw = (K / 64) * jbv * 2 * 64;
            DT_LAUNCH("ttm_m0_swizzle", m0_factor_swizzle, dim3((unsigned)std::min<int64_t>(ceil_div(nsw, 256), 1024)),
                      dim3(256), 0, fv.Ub16, (int)K, (int)J, jbv, Um);
            const int64_t ngroups = B / 32;
            // one workgroup per CU (the factor fills the LDS); PD: steps in flight per wave
            int cus = 0;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus < 1)
                cus = 256;
            dim3 grid((unsigned)std::min<int64_t>(cus, ceil_div(ngroups, 4)));
            const int nst = (int)(K / 64);
            const int pdv = nst % 8 == 0 ? 8 : (nst % 4 == 0 ? 4 : 2);
#define M0L(JBV, PDV)                                                                          \
    if (stream) {                                                                              \
        DT_LAUNCH("ttm_m0_bf16_stream", (ttm_m0_bf16_lds<JBV, PDV, true>), grid, dim3(256), 0, \
                  (const uint16_t *)A, Um, (float *)Y, (int)K, ngroups, (int)J);               \
    } else {                                                                                   \
        DT_LAUNCH("ttm_m0_bf16", (ttm_m0_bf16_lds<JBV, PDV, false>), grid, dim3(256), 0,       \
                  (const uint16_t *)A, Um, (float *)Y, (int)K, ngroups, (int)J);               \
    }                                                                                          \
    return DT_OK;
#define M0LP(JBV)                   \
    if (pdv == 8) { M0L(JBV, 8) }   \
    if (pdv == 4) { M0L(JBV, 4) }   \
    M0L(JBV, 2)
            if (jbv == 1) { M0LP(1) }
            if (jbv == 2) { M0LP(2) }
            M0LP(4)
#undef M0LP
#undef M0L
        }
        if (J <= 64 && F == 1 && (K % 8) == 0 && fv.Ub16) {
            dim3 grid((unsigned)ceil_div(B, 256));
#define M0B(JBV)                                                                          \
    if (stream) {                                                                         \
        DT_LAUNCH("ttm_m0_bf16_stream", (ttm_m0_bf16<JBV, true>), grid, dim3(256), 0,     \
                  (const uint16_t *)A, fv.Ub16, (float *)Y, (int)K, B, (int)J);           \
    } else {                                                                              \
        DT_LAUNCH("ttm_m0_bf16", (ttm_m0_bf16<JBV, false>), grid, dim3(256), 0,           \
                  (const uint16_t *)A, fv.Ub16, (float *)Y, (int)K, B, (int)J);           \
    }                                                                                     \
    return DT_OK;
            if (jbn == 1) { M0B(1) }
            if (jbn == 2) { M0B(2) }
            M0B(4)
#undef M0B
        }
        if (J <= 64 && F > 1 && (F % 8) == 0 && (K % 8) == 0 && fv.Ub16) {
            const int64_t tiles = ceil_div(F, 128);
            dim3 grid((unsigned)ceil_div(tiles * B, 4));
#define MKB(JBV)                                                                          \
    if (stream) {                                                                         \
        DT_LAUNCH("ttm_mk_bf16_stream", (ttm_mk_bf16<JBV, true>), grid, dim3(256), 0,     \
                  (const uint16_t *)A, fv.Ub16, (float *)Y, F, (int)K, B, (int)J, tiles); \
    } else {                                                                              \
        DT_LAUNCH("ttm_mk_bf16", (ttm_mk_bf16<JBV, false>), grid, dim3(256), 0,           \
                  (const uint16_t *)A, fv.Ub16, (float *)Y, F, (int)K, B, (int)J, tiles); \
    }                                                                                     \
    return DT_OK;
            if (jbn == 1) { MKB(1) }
            if (jbn == 2) { MKB(2) }
            MKB(4)
#undef MKB
        }
        DT_CHECK_ARG(fv.M32, "ttm: missing FP32 factor");
        DT_LAUNCH("ttm_generic_bf16", (ttm_generic<uint16_t, float, float, float>), gs_grid(total),
                  dim3(256), 0, (const uint16_t *)A, fv.M32, (float *)Y, F, K, B, J);
        return DT_OK;
    }
    return dt_fail(ctx, DT_EUNSUPPORTED, "ttm: unsupported input type %d", dtA);
}

static void mode_split(int ndim, const size_t *shape, int n, int64_t &F, int64_t &K, int64_t &B) {
    F = 1;
    B = 1;
    for (int i = 0; i < n; i++) F *= (int64_t)shape[i];
    K = (int64_t)shape[n];
    for (int i = n + 1; i < ndim; i++) B *= (int64_t)shape[i];
}

extern "C" int dt_ttm(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape,
                      const void *M, int dtM, size_t J, int n, void *Y, int dtY) {
    DT_CHECK_ARG(ctx && A && M && Y && shape, "dt_ttm: null argument");
    DT_CHECK_ARG(ndim >= 1 && ndim <= DT_MAX_NDIM && n >= 0 && n < ndim, "dt_ttm: bad mode %d of %d",
                 n, ndim);
    DT_CHECK_ARG(J >= 1, "dt_ttm: J must be positive");
    DT_CHECK_ARG(dtM == DT_F64 || dtM == DT_F32, "dt_ttm: factor must be F64 or F32");
    int64_t F, K, B;
    mode_split(ndim, shape, n, F, K, B);
    const int64_t mk = (int64_t)J * K;
    dt_factor_views fv = {};
    if (dtA == DT_F64) {
        if (dtM == DT_F64) {
            fv.M64 = (const double *)M;
        } else {
            double *m64 = (double *)dt_ws(ctx, WS_TTM_M, mk * sizeof(double));
            if (!m64) return DT_ENOMEM;
            DT_LAUNCH("factor_views", (copy_convert<double, float>), gs_grid(mk), dim3(256), 0, m64,
                      (const float *)M, mk);
            fv.M64 = m64;
        }
    } else {
        // workspace: M32 (J x K) | U32 (K x J) | Ub16 (K x J)
        char *ws = (char *)dt_ws(ctx, WS_TTM_M, mk * (4 + 4 + 2) + 512);
        if (!ws) return DT_ENOMEM;
        float *m32 = (float *)ws;
        float *u32 = (float *)(ws + ((mk * 4 + 255) & ~255LL));
        uint16_t *ub = (uint16_t *)((char *)u32 + ((mk * 4 + 255) & ~255LL));
        if (dtM == DT_F64) {
            DT_LAUNCH("factor_views", (copy_convert<float, double>), gs_grid(mk), dim3(256), 0, m32,
                      (const double *)M, mk);
        } else {
            DT_HIP(hipMemcpyAsync(m32, M, mk * 4, hipMemcpyDeviceToDevice, ctx->stream));
        }
        // M is J x K; U = M^T is K x J.
        DT_LAUNCH("factor_views", (transpose_convert<float, float>), gs_grid(mk), dim3(256), 0, u32,
                  m32, (int64_t)J, K);
        DT_LAUNCH("factor_views", (copy_convert<uint16_t, float>), gs_grid(mk), dim3(256), 0, ub, u32,
                  mk);
        fv.M32 = m32;
        fv.U32 = u32;
        fv.Ub16 = ub;
    }
    return dt_ttm_views(ctx, A, dtA, F, K, B, fv, (int64_t)J, Y, dtY, true);
}
