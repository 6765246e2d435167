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
__device__ __forceinline__ floatx4 ld_nt_f4(const float *p, bool nt) {
    return nt ? __builtin_nontemporal_load((const floatx4 *)p) : *(const floatx4 *)p;
}
__device__ __forceinline__ void wave_sync_wl() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int L2_YS = 68;  // LDS row stride (floats) of the J0 x 64 stage-1 block

// Per-type stage-1 operand traffic.  One "step" covers KSTEP consecutive i0
// of 64 columns i1: lane (c, g) holds 2 x 16 bytes of A for each of the four
// 16-column groups and 2 x 16 bytes of U0 for each 16-row group of j0.
template <bool BF16>
struct LeadTraits;
template <>
struct LeadTraits<false> {
    static constexpr int KSTEP = 32;
    typedef floatx4 vec;
    typedef float elem;
    static constexpr int LANE_K = 4;  // elements per lane per 16-byte load
};
template <>
struct LeadTraits<true> {
    static constexpr int KSTEP = 64;
    typedef u32x4 vec;
    typedef uint16_t elem;
    static constexpr int LANE_K = 8;
};

// Requirements (checked by the host): K1 % 64 == 0, K0 % 64 (FP32) or
// K0 % 128 (BF16) == 0, J0 <= 16 * JB0 <= 32, J1 <= 16 * JB1 <= 64.
//
// Every wave works alone on whole slabs b = wave id, wave id + #waves, ...
// A slab is processed as K1 / 64 units of 64 columns; a unit is K0 / KSTEP
// steps.  The loads run PD steps ahead of the MFMAs on one continuous
// cursor over (slab, unit, step), so the next unit's first loads are in
// flight while a unit's stage-1 block goes through LDS into stage 2, and no
// workgroup barrier is ever needed.
template <int JB0, int JB1, bool BF16>
__global__ __launch_bounds__(256) void ttm2_lead(const void *__restrict__ Av, const void *__restrict__ U0v,
                                                 const float *__restrict__ M1, float *__restrict__ T, int K0,
                                                 int K1, int J0, int J1, int64_t B) {
    typedef LeadTraits<BF16> Tr;
    typedef typename Tr::vec vec;
    typedef typename Tr::elem elem;
    constexpr int PD = 2;
    constexpr int KSTEP = Tr::KSTEP;
    __shared__ float Ysm[4][16 * JB0][L2_YS];
    // the stage-2 factor as M1s[i1][j1] when it fits (one fill per workgroup):
    // per-unit global loads would sit behind the in-flight tensor loads
    constexpr int M1_LDS = 8192;
    __shared__ float M1s[M1_LDS];
    const bool m1lds = (int64_t)K1 * 16 * JB1 <= M1_LDS;
    if (m1lds) {
        for (int e = threadIdx.x; e < K1 * 16 * JB1; e += 256) {
            const int i1 = e / (16 * JB1), j = e % (16 * JB1);
            M1s[e] = j < J1 ? M1[j + (int64_t)J1 * i1] : 0.f;
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = lane & 15;
    const int g = lane >> 4;
    const int64_t nw = (int64_t)gridDim.x * 4;
    const int64_t w0 = (int64_t)blockIdx.x * 4 + wave;
    if (w0 >= B) return;  // waves never synchronise with each other
    float(*Ys)[L2_YS] = Ysm[wave];
    const elem *A = (const elem *)Av;
    const elem *U0 = (const elem *)U0v;
    const int64_t slab = (int64_t)K0 * K1;
    const int units = K1 / 64;

    int jrow[JB0];
    bool jin[JB0];
#pragma unroll
    for (int jb = 0; jb < JB0; jb++) {
        const int j = 16 * jb + c;
        jrow[jb] = j < J0 ? j : J0 - 1;
        jin[jb] = j < J0;
    }
    int j1col[JB1];
    float j1mask[JB1];
#pragma unroll
    for (int jb = 0; jb < JB1; jb++) {
        const int j = 16 * jb + c;
        j1col[jb] = j < J1 ? j : J1 - 1;
        j1mask[jb] = j < J1 ? 1.f : 0.f;
    }
    // this lane's U0 row offsets (k added per step)
    const elem *Ul[JB0];
#pragma unroll
    for (int jb = 0; jb < JB0; jb++) Ul[jb] = U0 + Tr::LANE_K * g + (int64_t)K0 * jrow[jb];

    vec a[PD][4][2], u[PD][JB0][2];
    // load cursor: slab lb, unit lu, step offset lk
    int64_t lb = w0;
    int lu = 0, lk = 0;
    const elem *lbase = A + slab * lb + (int64_t)K0 * c + Tr::LANE_K * g;  // + K0 * 64 lu
    // past the last slab the loads repeat the slab before the end (never
    // used): every path issues the same loads, so the waits stay exact
    const elem *lclamp = A + slab * (B - 1) + (int64_t)K0 * c + Tr::LANE_K * g;
    auto issue = [&](int p) {
        {
            const elem *Ap = (lb < B ? lbase : lclamp) + (int64_t)K0 * 64 * lu + lk;
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int h = 0; h < 2; h++)
                    a[p][cb][h] = nt_load((const vec *)(Ap + (int64_t)K0 * 16 * cb + (KSTEP / 2) * h));
#pragma unroll
            for (int jb = 0; jb < JB0; jb++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    vec v = *(const vec *)(Ul[jb] + lk + (KSTEP / 2) * h);
                    if constexpr (BF16) {
                        const vec z = {0u, 0u, 0u, 0u};
                        u[p][jb][h] = jin[jb] ? v : z;
                    } else {
                        u[p][jb][h] = v * (jin[jb] ? 1.f : 0.f);
                    }
                }
        }
        lk += KSTEP;
        if (lk == K0) {
            lk = 0;
            if (++lu == units) {
                lu = 0;
                lb += nw;
                lbase += slab * nw;
            }
        }
    };
    floatx4 acc[JB0][4];
    auto step = [&](int p) {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int jb = 0; jb < JB0; jb++)
#pragma unroll
                for (int cb = 0; cb < 4; cb++) {
                    if constexpr (BF16) {
                        acc[jb][cb] = mfma_bf(u[p][jb][h], a[p][cb][h], acc[jb][cb]);
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; q++)
                            acc[jb][cb] = mfma4(u[p][jb][h][q], a[p][cb][h][q], acc[jb][cb]);
                    }
                }
    };

#pragma unroll
    for (int p = 0; p < PD; p++) issue(p);
    for (int64_t b = w0; b < B; b += nw) {
        floatx4 acc2[JB0][JB1];
#pragma unroll
        for (int i = 0; i < JB0; i++)
#pragma unroll
            for (int j = 0; j < JB1; j++) acc2[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
        for (int un = 0; un < units; un++) {
            const int cbase = 64 * un;
            // this unit's stage-2 factor values (from LDS when it holds M1)
            float m[16][JB1];
            if (!m1lds) {
#pragma unroll
                for (int q = 0; q < 16; q++)
#pragma unroll
                    for (int jb = 0; jb < JB1; jb++)
                        m[q][jb] = M1[j1col[jb] + (int64_t)J1 * (cbase + 4 * q + g)];
            }
#pragma unroll
            for (int jb = 0; jb < JB0; jb++)
#pragma unroll
                for (int cb = 0; cb < 4; cb++) acc[jb][cb] = floatx4{0.f, 0.f, 0.f, 0.f};
            for (int ks = 0; ks < K0; ks += KSTEP * PD) {
#pragma unroll
                for (int p = 0; p < PD; p++) {
                    step(p);
                    issue(p);
                }
            }
            // stage-1 block -> LDS as Ys[j0][i1 - cbase] (read back by this wave only)
#pragma unroll
            for (int jb = 0; jb < JB0; jb++)
#pragma unroll
                for (int cb = 0; cb < 4; cb++)
#pragma unroll
                    for (int v = 0; v < 4; v++) Ys[16 * jb + 4 * g + v][16 * cb + c] = acc[jb][cb][v];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // stage 2: acc2[j0][j1] += sum_i1 Ys[j0][i1] M1[j1, i1]
            if (m1lds) {
#pragma unroll
                for (int q = 0; q < 16; q++)
#pragma unroll
                    for (int jb = 0; jb < JB1; jb++) m[q][jb] = M1s[(cbase + 4 * q + g) * 16 * JB1 + 16 * jb + c];
            }
#pragma unroll
            for (int q = 0; q < 16; q++) {
                float y[JB0];
#pragma unroll
                for (int jb = 0; jb < JB0; jb++) y[jb] = Ys[16 * jb + c][4 * q + g];
#pragma unroll
                for (int i = 0; i < JB0; i++)
#pragma unroll
                    for (int j = 0; j < JB1; j++) acc2[i][j] = mfma4(y[i], m[q][j] * j1mask[j], acc2[i][j]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        // acc2[i][j][v] = T[j0 = 16 i + 4 g + v][j1 = 16 j + c]
        float *Tb = T + (int64_t)J0 * J1 * b;
#pragma unroll
        for (int i = 0; i < JB0; i++)
#pragma unroll
            for (int j = 0; j < JB1; j++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int j0 = 16 * i + 4 * g + v, j1 = 16 * j + c;
                    if (j0 < J0 && j1 < J1) Tb[j0 + (int64_t)J0 * j1] = acc2[i][j][v];
                }
    }
}

// ---- lead pair with whole-column loads ------------------------------------------
//   ttm2_lead_wl: the same product as ttm2_lead for FP32, K0 = 256, J0, J1 <= 16.
// Every load instruction reads one whole 1 KB column (64 lanes x 16 B, the
// pattern that streams at ~7 TB/s); a group of GC columns goes through a
// per-wave LDS tile (column stride 260 floats, so the MFMA-layout reads below
// hit 16 different bank quads) and is read back as lane (c, g) <- column c,
// i0 = 16 s + 4 g .. +3 for the stage-1 MFMAs (U0 operand in registers).  The
// next group's loads are issued as soon as the current group is in LDS.  The
// stage-1 block goes through a small LDS tile to the operand layout of stage
// 2, which runs against U1 held in LDS.
constexpr int WL_CS = 260;  // LDS column stride (floats)
constexpr int WL_GC = 16;   // columns per LDS group

template <int GC, int K1>  // GC = 16 columns per LDS group, K1 = 256
__global__ __launch_bounds__(256) void ttm2_lead_wl(const float *__restrict__ A, const float *__restrict__ U0,
                                                    const float *__restrict__ M1, float *__restrict__ T, int J0,
                                                    int J1, int64_t B) {
    constexpr int K0 = 256, NG = K1 / GC;
    __shared__ float Ac[4][GC * WL_CS];
    __shared__ float Ps[4][GC * 20];
    __shared__ float M1s[4096];  // M1s[i1 * 16 + j1], K1 <= 256
    // uniform wave index: the slab loop stays scalar (exact waits across it)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = lane & 15, g = lane >> 4;
    for (int e = threadIdx.x; e < K1 * 16; e += 256) {
        const int i1 = e >> 4, j = e & 15;
        M1s[e] = j < J1 ? M1[j + (int64_t)J1 * i1] : 0.f;
    }
    __syncthreads();
    const int64_t nw = (int64_t)gridDim.x * 4;
    const int64_t w0 = (int64_t)blockIdx.x * 4 + wave;
    if (w0 >= B) return;  // no workgroup barrier after this point
    float *Aw = Ac[wave];
    float *Pw = Ps[wave];
    // stage-1 operand U0[16 s + 4 g + t, j0 = c], masked past J0 (U0: K0 x J0, i0 fastest)
    float u0[16][4];
    const bool jin = c < J0;
    const int jc = jin ? c : J0 - 1;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const floatx4 v = *(const floatx4 *)(U0 + 16 * q + 4 * g + (int64_t)K0 * jc);
#pragma unroll
        for (int t = 0; t < 4; t++) u0[q][t] = jin ? v[t] : 0.f;
    }
    constexpr int groups = NG;
    const int64_t slab = (int64_t)K0 * K1;
    // load cursor: slab lb, group lgr (past the last slab: reload the last one)
    int64_t lb = w0;
    int lgr = 0;
    // two register sets alternate (groups per slab is even), so the loads of
    // the next two groups are in flight while a group is computed
    floatx4 vA[GC], vB[GC];
    auto issue = [&](floatx4 (&v)[GC]) {
        const int64_t bb = lb < B ? lb : B - 1;
        const float *p = A + slab * bb + (int64_t)K0 * GC * lgr + 4 * lane;
#pragma unroll
        for (int q = 0; q < GC; q++) v[q] = ld_nt_f4(p + (int64_t)K0 * q, true);
        if (++lgr == groups) {
            lgr = 0;
            lb += nw;
        }
    };
    issue(vA);
    issue(vB);
    // per slab: NG = K1 / GC groups, fully unrolled in pairs (static register
    // sets and a static load sequence, so the waits stay exact)
    for (int64_t b = w0; b < B; b += nw) {
        floatx4 acc2 = {0.f, 0.f, 0.f, 0.f};
        auto group = [&](int gr, floatx4 (&v)[GC]) {
            wave_sync_wl();  // the previous group's LDS reads are done
#pragma unroll
            for (int q = 0; q < GC; q++) *(floatx4 *)(Aw + q * WL_CS + 4 * lane) = v[q];
            issue(v);
            wave_sync_wl();
            // stage 1 over this group's 16 columns
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const floatx4 a = *(const floatx4 *)(Aw + c * WL_CS + 16 * q + 4 * g);
#pragma unroll
                for (int t = 0; t < 4; t++) acc = mfma4(u0[q][t], a[t], acc);
            }
            // acc[v] = P[j0 = 4g + v][column c] -> Ps[column][j0]
            *(floatx4 *)(Pw + c * 20 + 4 * g) = acc;
            wave_sync_wl();
            // stage 2: acc2[j0][j1] += sum over the group's columns P[j0][col] U1[col][j1]
#pragma unroll
            for (int kk = 0; kk < GC / 4; kk++) {
                const float pa = Pw[(4 * kk + g) * 20 + c];
                const float mb = M1s[(GC * gr + 4 * kk + g) * 16 + c];
                acc2 = mfma4(pa, mb, acc2);
            }
        };
#pragma unroll
        for (int gr = 0; gr < NG; gr += 2) {
            group(gr, vA);
            group(gr + 1, vB);
        }
        // acc2[v] = T[j0 = 4g + v][j1 = c]
        float *Tb = T + (int64_t)J0 * J1 * b;
        if (c < J1)
#pragma unroll
            for (int vv = 0; vv < 4; vv++) {
                const int j0 = 4 * g + vv;
                if (j0 < J0) Tb[j0 + (int64_t)J0 * c] = acc2[vv];
            }
    }
}

// ---- trailing pair ------------------------------------------------------------
//   ttm2_trail:  T[f, j2, j3] = sum_{i2, i3} U2[i2, j2] U3[i3, j3] A[f, i2, i3]
// (the last two modes of a tensor viewed as F x K2 x K3, FP32).  A group of
// four waves owns 64 consecutive f, i.e. 256-byte pieces of every (i2, i3)
// column (a workgroup holds FG such groups on adjacent f, which the stage-2
// barrier keeps in step), and wave w of a group streams i2 in
// [w K2/4, (w + 1) K2/4) for every i3 on one continuous load cursor.  Stage 1 (per i3 and wave) is the
// ttm_mk_f32 MFMA form: A operand U2ᵀ (lane (c, g): M2[c, i2 + g]), B operand
// a 16-byte f-contiguous piece of A (lane (c, g): A[f0 + 4c + t, i2 + g, i3],
// t = 0..3 in four MFMAs), giving the wave's partial P_w[f0 + 4c + t, 4g + v].
// The four partials go to LDS (double-buffered, rows j2 * 64 + f); stage 2 is
// one MFMA per 16-row tile and i3 with K = the four waves (lane (c, g) reads
// P_g[row c]) and B = U3[i3, j3 = c] in every group, so the MFMA also adds the
// partials.  T (64 x J2 x J3 per workgroup) stays in registers until the end:
// the pass reads A once and writes |A| J2 J3 / (K2 K3) floats.
constexpr int TR_PSTRIDE = 1024 + 16;  // floats per wave partial (the pad staggers banks)

template <int KQ, bool NT, int FG>
__global__ __launch_bounds__(256 * FG) void ttm2_trail(const float *__restrict__ A, const float *__restrict__ M2,
                                                       const float *__restrict__ M3, float *__restrict__ T, int64_t F,
                                                       int K2, int K3, int J2, int J3) {
    // KQ = K2 / 16 load steps (4 i2 each) per i3 and wave; the ring holds one
    // whole i3 of steps, so the next i3's loads are in flight through the
    // stage-2 barrier.  FG groups of four waves own FG adjacent 64-f blocks,
    // so the workgroup reads 256 FG contiguous bytes of every column in step.
    __shared__ float Ps[2][FG * 4 * TR_PSTRIDE];
    const int lane = threadIdx.x & 63, fg = threadIdx.x >> 8, wave = (threadIdx.x >> 6) & 3;
    const int c = lane & 15, g = lane >> 4;
    const int64_t f0 = (int64_t)blockIdx.x * 64 * FG + 64 * fg;
    const int i2w = wave * (K2 / 4);
    const float um = c < J2 ? 1.f : 0.f;
    const int cj2 = c < J2 ? c : J2 - 1;
    const float m3m = c < J3 ? 1.f : 0.f;
    const int cj3 = c < J3 ? c : J3 - 1;
    // the U2 operand of every step of this wave (its i2 range is fixed)
    float u[KQ];
#pragma unroll
    for (int q = 0; q < KQ; q++) u[q] = M2[cj2 + (int64_t)J2 * (i2w + 4 * q + g)] * um;

    const float *Ab = A + f0 + 4 * c + F * (int64_t)(i2w + g);
    // two half rings: while the MFMAs run over one half of i3's steps, the
    // other half's loads for i3 + 1 are already in flight (scheduling barriers
    // keep the compiler from sinking them below the MFMAs)
    constexpr int HQ = KQ / 2;
    floatx4 r0[HQ], r1[HQ];
    auto ld = [&](int q, int i3) { return ld_nt_f4(Ab + F * ((int64_t)4 * q + (int64_t)K2 * i3), NT); };
#pragma unroll
    for (int q = 0; q < HQ; q++) r0[q] = ld(q, 0);
#pragma unroll
    for (int q = 0; q < HQ; q++) r1[q] = ld(HQ + q, 0);
    float b3n = M3[cj3] * m3m;
    floatx4 acc2[16];
#pragma unroll
    for (int t = 0; t < 16; t++) acc2[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int i3 = 0; i3 < K3; i3++) {
        floatx4 acc[4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
        const float b3 = b3n;
        // unconditional (the last i3 reloads itself) so that every path issues
        // the same loads and the waits stay exact
        const int in3 = i3 + 1 < K3 ? i3 + 1 : i3;
#pragma unroll
        for (int q = 0; q < HQ; q++) {
            acc[0] = mfma4(u[q], r0[q][0], acc[0]);
            acc[1] = mfma4(u[q], r0[q][1], acc[1]);
            acc[2] = mfma4(u[q], r0[q][2], acc[2]);
            acc[3] = mfma4(u[q], r0[q][3], acc[3]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < HQ; q++) r0[q] = ld(q, in3);
        b3n = M3[cj3 + (int64_t)J3 * in3] * m3m;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < HQ; q++) {
            acc[0] = mfma4(u[HQ + q], r1[q][0], acc[0]);
            acc[1] = mfma4(u[HQ + q], r1[q][1], acc[1]);
            acc[2] = mfma4(u[HQ + q], r1[q][2], acc[2]);
            acc[3] = mfma4(u[HQ + q], r1[q][3], acc[3]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < HQ; q++) r1[q] = ld(HQ + q, in3);
        __builtin_amdgcn_sched_barrier(0);
        // partial of this wave: acc[t][v] = P_w[f0 + 4c + t, j2 = 4g + v]; row j2 * 64 + f
        float *P = Ps[i3 & 1] + (4 * fg + wave) * TR_PSTRIDE;
#pragma unroll
        for (int v = 0; v < 4; v++)
            *(floatx4 *)(P + (4 * g + v) * 64 + 4 * c) = floatx4{acc[0][v], acc[1][v], acc[2][v], acc[3][v]};
        __syncthreads();
        // stage 2: tiles 16 wave .. 16 wave + 15 of the 1024 rows
        const float *Pg = Ps[i3 & 1] + (4 * fg + g) * TR_PSTRIDE;
#pragma unroll
        for (int t = 0; t < 16; t++) acc2[t] = mfma4(Pg[(16 * wave + t) * 16 + c], b3, acc2[t]);
    }
    // acc2[t][v]: row r = 16 (16 wave + t) + 4g + v -> j2 = r / 64, f = f0 + r % 64; column j3 = c
    if (c < J3) {
#pragma unroll
        for (int t = 0; t < 16; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = 16 * (16 * wave + t) + 4 * g + v;
                const int j2 = r >> 6;
                if (j2 < J2) T[f0 + (r & 63) + F * (j2 + (int64_t)J2 * c)] = acc2[t][v];
            }
    }
}

static bool ttm2_trail_fits(int dtA, int64_t F, int64_t K2, int64_t K3, int64_t J2, int64_t J3) {
    return dtA == DT_F32 && F >= 64 && F % 64 == 0 && (K2 == 128 || K2 == 256) && K3 >= 1 &&
           K3 <= 0x7fffffff && J2 >= 1 && J2 <= 16 && J3 >= 1 && J3 <= 16 && F / 64 <= 0x7fffffff;
}

// T (F x J2 x J3) = A (F x K2 x K3) x_1 M2 x_2 M3 with M2 = J2 x K2 and
// M3 = J3 x K3 (FP32, j fastest).  DT_EUNSUPPORTED when the shape does not
// fit (the caller then runs two single products).
int dt_ttm2_trail(dt_ctx *ctx, const float *A, int dtA, int64_t F, int64_t K2, int64_t K3, const float *M2,
                  const float *M3, int64_t J2, int64_t J3, float *T) {
    if (!ttm2_trail_fits(dtA, F, K2, K3, J2, J3)) return DT_EUNSUPPORTED;
    // two 64-f blocks per workgroup when F allows: 512 contiguous bytes of
    // every column per step (2.86-2.88 ms against 2.93-2.99 ms at 256^4)
    if (K2 == 256 && F % 128 == 0) {
        DT_LAUNCH("ttm2_trail_f32_stream", (ttm2_trail<16, true, 2>), dim3((unsigned)(F / 128)), dim3(512), 0, A, M2,
                  M3, T, F, (int)K2, (int)K3, (int)J2, (int)J3);
    } else if (K2 == 256) {
        DT_LAUNCH("ttm2_trail_f32_stream", (ttm2_trail<16, true, 1>), dim3((unsigned)(F / 64)), dim3(256), 0, A, M2, M3,
                  T, F, (int)K2, (int)K3, (int)J2, (int)J3);
    } else {
        DT_LAUNCH("ttm2_trail_f32_stream", (ttm2_trail<8, true, 1>), dim3((unsigned)(F / 64)), dim3(256), 0, A, M2, M3,
                  T, F, (int)K2, (int)K3, (int)J2, (int)J3);
    }
    return DT_OK;
}

// waves resident on the context's device for a kernel at 256 threads
template <typename K>
static int64_t resident_waves(dt_ctx *ctx, K kernel) {
    return dt_resident_wgs(ctx, kernel, 256) * 4;
}

static bool ttm2_lead_fits(int dtA, int64_t K0, int64_t K1, int64_t B, int64_t J0, int64_t J1) {
    if (dtA != DT_F32 && dtA != DT_BF16) return false;
    return K0 > 0 && K1 > 0 && K1 % 64 == 0 && K0 % (dtA == DT_BF16 ? 128 : 64) == 0 && J0 >= 1 &&
           J0 <= 32 && J1 >= 1 && J1 <= 64 && B >= 1 && K0 <= 0x7fffffff && K1 <= 0x7fffffff;
}

// T (J0 x J1 x B) = A (K0 x K1 x B) x_0 U0ᵀ x_1 U1ᵀ.  U0 given as K0 x J0
// (FP32, or bf16 for a BF16 tensor), U1 as M1 = J1 x K1 FP32.  Returns
// DT_EUNSUPPORTED when the shape does not fit the kernel (the caller then
// runs two single products).
int dt_ttm2_lead(dt_ctx *ctx, const void *A, int dtA, int64_t K0, int64_t K1, int64_t B,
                 const void *U0, const float *M1, int64_t J0, int64_t J1, float *T) {
    if (!ttm2_lead_fits(dtA, K0, K1, B, J0, J1)) return DT_EUNSUPPORTED;
    if (dtA == DT_F32 && K0 == 256 && K1 == 256 && J0 <= 16 && J1 <= 16) {
        const int64_t waves = resident_waves(ctx, ttm2_lead_wl<WL_GC, 256>);
        if (waves <= 0) return dt_fail(ctx, DT_EHIP, "dt_ttm2_lead: occupancy query failed");
        const int64_t nwave = std::min<int64_t>(B, waves);
        DT_LAUNCH("ttm2_lead_f32_stream", (ttm2_lead_wl<WL_GC, 256>), dim3((unsigned)((nwave + 3) / 4)), dim3(256), 0,
                  (const float *)A, (const float *)U0, M1, T, (int)J0, (int)J1, B);
        return DT_OK;
    }
    const bool bf = dtA == DT_BF16;
    const int jb0 = J0 <= 16 ? 1 : 2;
    const int jb1 = J1 <= 16 ? 1 : (J1 <= 32 ? 2 : 4);
#define L2K(A0, A1)                                                                                   \
    if (jb0 == A0 && jb1 == A1) {                                                                     \
        auto kern = bf ? ttm2_lead<A0, A1, true> : ttm2_lead<A0, A1, false>;                           \
        const int64_t waves = resident_waves(ctx, kern);                                              \
        if (waves <= 0) return dt_fail(ctx, DT_EHIP, "dt_ttm2_lead: occupancy query failed");        \
        const int64_t nwave = std::min<int64_t>(B, waves);                                            \
        dim3 grid((unsigned)((nwave + 3) / 4));                                                       \
        DT_LAUNCH(bf ? "ttm2_lead_bf16_stream" : "ttm2_lead_f32_stream", kern, grid, dim3(256), 0, A,  \
                  U0, M1, T, (int)K0, (int)K1, (int)J0, (int)J1, B);                                  \
        return DT_OK;                                                                                 \
    }
    L2K(1, 1) L2K(1, 2) L2K(1, 4) L2K(2, 1) L2K(2, 2) L2K(2, 4)
#undef L2K
    return DT_EUNSUPPORTED;
}

// Factor views for the fused pair from J x K column-major factors (reference
// convention Y = A x_n M): U0 = M0ᵀ (K0 x J0, FP32 or bf16), M1 FP32 copy.
template <typename TS, typename TD>
__global__ void pair_view(TD *__restrict__ dst, const TS *__restrict__ src, int64_t J, int64_t K, bool transpose) {
    const int64_t total = J * K;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e % J, k = e / J;  // src[j + J k]
        const float v = (float)src[e];
        const int64_t o = transpose ? k + K * j : e;
        if constexpr (sizeof(TD) == 2)
            dst[o] = f32_to_bf16(v);
        else
            dst[o] = v;
    }
}

static inline dim3 pv_grid(int64_t n) { return dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)); }

// Y = A x_{idx[0]} M[0] x_{idx[1]} M[1] ... (Function::ttmc).  When the chain
// starts with modes {0, 1} of an FP32/BF16 tensor, those two products run as
// one pass (ttm2_lead); the rest are single products through dt_ttm.
extern "C" int dt_ttmc(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape,
                       const void *const *M, int dtM, const size_t *J, const int *idx, int nm, void *Y,
                       int dtY) {
    DT_CHECK_ARG(ctx && A && M && J && idx && Y && shape, "dt_ttmc: null argument");
    DT_CHECK_ARG(ndim >= 1 && ndim <= DT_MAX_NDIM, "dt_ttmc: bad ndim %d", ndim);
    DT_CHECK_ARG(nm >= 1, "dt_ttmc: need at least one factor");
    DT_CHECK_ARG(dtA == DT_F64 || dtA == DT_F32 || dtA == DT_BF16, "dt_ttmc: bad tensor type %d", dtA);
    DT_CHECK_ARG(dtM == DT_F64 || dtM == DT_F32, "dt_ttmc: factor must be F64 or F32");
    DT_CHECK_ARG(dtY == (dtA == DT_F64 ? DT_F64 : DT_F32), "dt_ttmc: output type %d does not match input %d",
                 dtY, dtA);
    for (int s = 0; s < nm; s++) {
        DT_CHECK_ARG(idx[s] >= 0 && idx[s] < ndim, "dt_ttmc: bad mode %d", idx[s]);
        DT_CHECK_ARG(M[s] && J[s] >= 1, "dt_ttmc: factor %d missing or empty", s);
    }
    size_t sh[DT_MAX_NDIM];
    for (int i = 0; i < ndim; i++) sh[i] = shape[i];
    const int idt = dtA == DT_F64 ? DT_F64 : DT_F32;  // intermediate type
    const size_t isz = idt == DT_F64 ? 8 : 4;
    const void *cur = A;
    int curdt = dtA;
    int s = 0;
    auto scratch = [&](int step, size_t elems) -> void * {
        return dt_ws(ctx, (step & 1) ? WS_TTMC1 : WS_TTMC0, elems * isz);
    };
    const bool pair = nm >= 2 && ndim >= 2 && ((idx[0] == 0 && idx[1] == 1) || (idx[0] == 1 && idx[1] == 0));
    const int p0 = idx[0] == 0 ? 0 : 1;  // position of the mode-0 factor
    int64_t B = 1;
    for (int i = 2; i < ndim; i++) B *= (int64_t)sh[i];
    if (pair && ttm2_lead_fits(dtA, (int64_t)sh[0], (int64_t)sh[1], B, (int64_t)J[p0], (int64_t)J[1 - p0])) {
        const int64_t K0 = (int64_t)sh[0], K1 = (int64_t)sh[1], J0 = (int64_t)J[p0], J1 = (int64_t)J[1 - p0];
        const bool last = nm == 2;
        // the pair counts as step 1: the next product reads slot 1, writes slot 0
        float *T = last ? (float *)Y : (float *)scratch(1, (size_t)(J0 * J1 * B));
        const bool bf = dtA == DT_BF16;
        const size_t u0b = ((size_t)(K0 * J0) * (bf ? 2 : 4) + 255) & ~(size_t)255;
        char *ws = (char *)dt_ws(ctx, WS_TTMC_M, u0b + (size_t)(J1 * K1) * 4);
        if (!T || !ws) return DT_ENOMEM;
        void *U0 = ws;
        float *M1 = (float *)(ws + u0b);
        const void *M0s = M[p0], *M1s = M[1 - p0];
        // U0 = M0ᵀ (K0 x J0) in the tensor's operand type, M1 as FP32 J1 x K1
        if (dtM == DT_F64 && bf) {
            DT_LAUNCH("factor_views", (pair_view<double, uint16_t>), pv_grid(K0 * J0), dim3(256), 0,
                      (uint16_t *)U0, (const double *)M0s, J0, K0, true);
        } else if (dtM == DT_F64) {
            DT_LAUNCH("factor_views", (pair_view<double, float>), pv_grid(K0 * J0), dim3(256), 0,
                      (float *)U0, (const double *)M0s, J0, K0, true);
        } else if (bf) {
            DT_LAUNCH("factor_views", (pair_view<float, uint16_t>), pv_grid(K0 * J0), dim3(256), 0,
                      (uint16_t *)U0, (const float *)M0s, J0, K0, true);
        } else {
            DT_LAUNCH("factor_views", (pair_view<float, float>), pv_grid(K0 * J0), dim3(256), 0,
                      (float *)U0, (const float *)M0s, J0, K0, true);
        }
        if (dtM == DT_F64) {
            DT_LAUNCH("factor_views", (pair_view<double, float>), pv_grid(K1 * J1), dim3(256), 0, M1,
                      (const double *)M1s, J1, K1, false);
        } else {
            DT_LAUNCH("factor_views", (pair_view<float, float>), pv_grid(K1 * J1), dim3(256), 0, M1,
                      (const float *)M1s, J1, K1, false);
        }
        DT_TRY(dt_ttm2_lead(ctx, A, dtA, K0, K1, B, U0, M1, J0, J1, T));
        if (last) return DT_OK;
        sh[0] = (size_t)J0;
        sh[1] = (size_t)J1;
        cur = T;
        curdt = DT_F32;
        s = 2;
    }
    for (; s < nm; s++) {
        const bool last = s == nm - 1;
        size_t out = 1;
        for (int i = 0; i < ndim; i++) out *= i == idx[s] ? J[s] : sh[i];
        void *dst = last ? Y : scratch(s, out);
        if (!dst) return DT_ENOMEM;
        DT_TRY(dt_ttm(ctx, cur, curdt, ndim, sh, M[s], dtM, J[s], idx[s], dst, last ? dtY : idt));
        sh[idx[s]] = J[s];
        cur = dst;
        curdt = last ? dtY : idt;
    }
    return DT_OK;
}
