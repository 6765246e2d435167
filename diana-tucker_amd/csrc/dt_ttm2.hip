// Two mode products in one pass over the tensor.
//
// In the reference each Function::ttm call (template/function/tensor.tpp:272-335)
// materialises its full-size result; the dimension-tree schedule of
// dt_hooi.hip would still write the output of the first product of each
// phase (|A| R / I bytes) and read it back for the second.  On MI355X a
// stream that mixes reads with 1/16 as many writes runs at ~5.6 TB/s against
// ~7 TB/s for a read-only stream, so this kernel keeps the first product in
// registers and LDS and writes only the doubly contracted result:
//
//   ttm2_lead:  T[j0, j1, b] = sum_{i0, i1} U0[i0, j0] U1[i1, j1] A[i0, i1, b]
//
// One workgroup (4 waves) owns one contiguous I0 x I1 slab of A (fixed b).
// Stage 1 is the ttm_m0 MFMA loop (each wave: 64 columns i1 at a time, i0
// streamed in 16-byte k-contiguous slices straight into the MFMA operand,
// a register ring of PD steps ahead); the J0 x 64 stage-1 block goes through
// LDS to reach the operand layout of stage 2, an MFMA over i1 against the
// J1 x I1 factor; the four waves' J0 x J1 partials are summed in LDS.
#include "dt_internal.hpp"

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mfma_bf(u32x4 a, u32x4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
template <typename T>
__device__ __forceinline__ T nt_load(const T *p) {
    return __builtin_nontemporal_load(p);
}

constexpr int L2_YS = 68;  // LDS row stride (floats) of the J0 x 64 stage-1 block

// Stage 1 for one 64-column block of the slab, FP32 input.
// Ab: slab base + K0 * (column block start); returns acc[jb][cb] with
// D[j0 = 16 jb + 4g + v][i1 = 16 cb + c].
template <int JB>
__device__ __forceinline__ void lead_stage1_f32(const float *__restrict__ Ab, const float *__restrict__ U0,
                                                int K0, int J0, floatx4 (&acc)[JB][4], int c, int g) {
    constexpr int PD = 2;
    int jrow[JB];
    float jmask[JB];
#pragma unroll
    for (int jb = 0; jb < JB; jb++) {
        const int j = 16 * jb + c;
        jrow[jb] = j < J0 ? j : J0 - 1;
        jmask[jb] = j < J0 ? 1.f : 0.f;
    }
    const float *Ap = Ab + (int64_t)K0 * c + 4 * g;
    floatx4 a[PD][4][2], u[PD][JB][2];
    auto load = [&](int p, int k0) {
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int h = 0; h < 2; h++)
                a[p][cb][h] = nt_load((const floatx4 *)(Ap + (int64_t)K0 * 16 * cb + k0 + 16 * h));
#pragma unroll
        for (int jb = 0; jb < JB; jb++)
#pragma unroll
            for (int h = 0; h < 2; h++)
                u[p][jb][h] = *(const floatx4 *)(U0 + k0 + 16 * h + 4 * g + (int64_t)K0 * jrow[jb]) * jmask[jb];
    };
    auto step = [&](int p) {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int jb = 0; jb < JB; jb++)
#pragma unroll
                for (int cb = 0; cb < 4; cb++) {
                    acc[jb][cb] = mfma4(u[p][jb][h][0], a[p][cb][h][0], acc[jb][cb]);
                    acc[jb][cb] = mfma4(u[p][jb][h][1], a[p][cb][h][1], acc[jb][cb]);
                    acc[jb][cb] = mfma4(u[p][jb][h][2], a[p][cb][h][2], acc[jb][cb]);
                    acc[jb][cb] = mfma4(u[p][jb][h][3], a[p][cb][h][3], acc[jb][cb]);
                }
    };
#pragma unroll
    for (int p = 0; p < PD; p++) load(p, 32 * p);
    int kb = 0;
    for (; kb + 32 * PD < K0; kb += 32 * PD) {
#pragma unroll
        for (int p = 0; p < PD; p++) {
            step(p);
            load(p, kb + 32 * (p + PD));
        }
    }
#pragma unroll
    for (int p = 0; p < PD; p++) step(p);
}

// Stage 1, BF16 input (v_mfma_f32_16x16x32_bf16), U0 in bf16 K0 x J0.
template <int JB>
__device__ __forceinline__ void lead_stage1_bf16(const uint16_t *__restrict__ Ab,
                                                 const uint16_t *__restrict__ U0, int K0, int J0,
                                                 floatx4 (&acc)[JB][4], int c, int g) {
    constexpr int PD = 2;
    int jrow[JB];
    bool jin[JB];
#pragma unroll
    for (int jb = 0; jb < JB; jb++) {
        const int j = 16 * jb + c;
        jrow[jb] = j < J0 ? j : J0 - 1;
        jin[jb] = j < J0;
    }
    const u32x4 zero4 = {0u, 0u, 0u, 0u};
    const uint16_t *Ap = Ab + (int64_t)K0 * c + 8 * g;
    u32x4 a[PD][4][2], u[PD][JB][2];
    auto load = [&](int p, int k0) {
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int h = 0; h < 2; h++)
                a[p][cb][h] = nt_load((const u32x4 *)(Ap + (int64_t)K0 * 16 * cb + k0 + 32 * h));
#pragma unroll
        for (int jb = 0; jb < JB; jb++)
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const u32x4 v = *(const u32x4 *)(U0 + k0 + 32 * h + 8 * g + (int64_t)K0 * jrow[jb]);
                u[p][jb][h] = jin[jb] ? v : zero4;
            }
    };
    auto step = [&](int p) {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int jb = 0; jb < JB; jb++)
#pragma unroll
                for (int cb = 0; cb < 4; cb++) acc[jb][cb] = mfma_bf(u[p][jb][h], a[p][cb][h], acc[jb][cb]);
    };
#pragma unroll
    for (int p = 0; p < PD; p++) load(p, 64 * p);
    int kb = 0;
    for (; kb + 64 * PD < K0; kb += 64 * PD) {
#pragma unroll
        for (int p = 0; p < PD; p++) {
            step(p);
            load(p, kb + 64 * (p + PD));
        }
    }
#pragma unroll
    for (int p = 0; p < PD; p++) step(p);
}

// Requirements (checked by the host): K1 % 64 == 0, K0 % 64 (FP32) or
// K0 % 128 (BF16) == 0, J0 <= 16 * JB0 <= 32, J1 <= 16 * JB1 <= 64.
template <int JB0, int JB1, bool BF16>
__global__ __launch_bounds__(256) void ttm2_lead(const void *__restrict__ Av, const void *__restrict__ U0v,
                                                 const float *__restrict__ M1, float *__restrict__ T,
                                                 int K0, int K1, int J0, int J1) {
    // stage-1 blocks per wave: [4][16 JB0][L2_YS]; reused for the final
    // cross-wave reduction [4][16 JB0 * 16 JB1] (fits: JB1 <= 4 < L2_YS / 16)
    __shared__ float smem[4 * 16 * JB0 * L2_YS];
    float (*Ys)[16 * JB0][L2_YS] = (float (*)[16 * JB0][L2_YS])smem;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c = lane & 15;
    const int g = lane >> 4;
    const int64_t b = blockIdx.x;
    const int64_t slab = (int64_t)K0 * K1;

    floatx4 acc2[JB0][JB1];
#pragma unroll
    for (int i = 0; i < JB0; i++)
#pragma unroll
        for (int j = 0; j < JB1; j++) acc2[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    int j1col[JB1];
    float j1mask[JB1];
#pragma unroll
    for (int jb = 0; jb < JB1; jb++) {
        const int j = 16 * jb + c;
        j1col[jb] = j < J1 ? j : J1 - 1;
        j1mask[jb] = j < J1 ? 1.f : 0.f;
    }

    for (int cbase = 64 * wave; cbase < K1; cbase += 256) {
        floatx4 acc[JB0][4];
#pragma unroll
        for (int jb = 0; jb < JB0; jb++)
#pragma unroll
            for (int cb = 0; cb < 4; cb++) acc[jb][cb] = floatx4{0.f, 0.f, 0.f, 0.f};
        if (BF16)
            lead_stage1_bf16<JB0>((const uint16_t *)Av + slab * b + (int64_t)K0 * cbase,
                                  (const uint16_t *)U0v, K0, J0, acc, c, g);
        else
            lead_stage1_f32<JB0>((const float *)Av + slab * b + (int64_t)K0 * cbase, (const float *)U0v,
                                 K0, J0, acc, c, g);
        // stage-1 block -> LDS as Ys[j0][i1 - cbase]
#pragma unroll
        for (int jb = 0; jb < JB0; jb++)
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int v = 0; v < 4; v++) Ys[wave][16 * jb + 4 * g + v][16 * cb + c] = acc[jb][cb][v];
        // the block is read back by other lanes of the same wave only
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // stage 2: acc2[j0][j1] += sum_i1 Ys[j0][i1] M1[j1, i1]
#pragma unroll 4
        for (int ks = 0; ks < 64; ks += 4) {
            float a[JB0], m[JB1];
#pragma unroll
            for (int jb = 0; jb < JB0; jb++) a[jb] = Ys[wave][16 * jb + c][ks + g];
#pragma unroll
            for (int jb = 0; jb < JB1; jb++) m[jb] = M1[j1col[jb] + (int64_t)J1 * (cbase + ks + g)] * j1mask[jb];
#pragma unroll
            for (int i = 0; i < JB0; i++)
#pragma unroll
                for (int j = 0; j < JB1; j++) acc2[i][j] = mfma4(a[i], m[j], acc2[i][j]);
        }
        // all lanes' reads of this block precede the next block's writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();  // every wave is done with its Ys before the space is reused
    float (*red)[16 * JB0 * 16 * JB1] = (float (*)[16 * JB0 * 16 * JB1])smem;
    // reduce the four waves' partials: red[w][j0 + J0p * j1]
    constexpr int J0p = 16 * JB0;
#pragma unroll
    for (int i = 0; i < JB0; i++)
#pragma unroll
        for (int j = 0; j < JB1; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) red[wave][(16 * i + 4 * g + v) + J0p * (16 * j + c)] = acc2[i][j][v];
    __syncthreads();
    for (int e = threadIdx.x; e < J0 * J1; e += 256) {
        const int j0 = e % J0, j1 = e / J0;
        const int o = j0 + J0p * j1;
        T[e + (int64_t)J0 * J1 * b] = (red[0][o] + red[1][o]) + (red[2][o] + red[3][o]);
    }
}

// T (J0 x J1 x B) = A (K0 x K1 x B) x_0 U0ᵀ x_1 U1ᵀ.  U0 given as K0 x J0
// (FP32, or bf16 for a BF16 tensor), U1 as M1 = J1 x K1 FP32.  Returns
// DT_EUNSUPPORTED when the shape does not fit the kernel (the caller then
// runs two single products).
int dt_ttm2_lead(dt_ctx *ctx, const void *A, int dtA, int64_t K0, int64_t K1, int64_t B,
                 const void *U0, const float *M1, int64_t J0, int64_t J1, float *T) {
    const bool bf = dtA == DT_BF16;
    if (!(dtA == DT_F32 || bf)) return DT_EUNSUPPORTED;
    if (K1 % 64 || K0 % (bf ? 128 : 64) || J0 > 32 || J1 > 64 || B < 1 || B > 0x7fffffff)
        return DT_EUNSUPPORTED;
    const int jb0 = J0 <= 16 ? 1 : 2;
    const int jb1 = J1 <= 16 ? 1 : (J1 <= 32 ? 2 : 4);
    dim3 grid((unsigned)B);
#define L2K(A0, A1)                                                                                   \
    if (jb0 == A0 && jb1 == A1) {                                                                     \
        if (bf) {                                                                                     \
            DT_LAUNCH("ttm2_lead_bf16_stream", (ttm2_lead<A0, A1, true>), grid, dim3(256), 0, A, U0, M1, T, \
                      (int)K0, (int)K1, (int)J0, (int)J1);                                            \
        } else {                                                                                      \
            DT_LAUNCH("ttm2_lead_f32_stream", (ttm2_lead<A0, A1, false>), grid, dim3(256), 0, A, U0, M1, T, \
                      (int)K0, (int)K1, (int)J0, (int)J1);                                            \
        }                                                                                             \
        return DT_OK;                                                                                 \
    }
    L2K(1, 1) L2K(1, 2) L2K(1, 4) L2K(2, 1) L2K(2, 2) L2K(2, 4)
#undef L2K
    return DT_EUNSUPPORTED;
}
