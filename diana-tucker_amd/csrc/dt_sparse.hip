// Sparse HOOI: Zipf COO generator with duplicate merging, the SpTensor
// mode sorts, the TTMc and the one-step-per-mode sparse HOOI.
//
// Reference path (spmain.cpp:55 -> Algorithm::SpTucker::Sp_HOOI_ALS,
// template/algorithm/sp_tucker/sp_hooi_als.tpp:67-118): per sweep and mode n,
// Sp_TTMC_ALS_ (:48-65) forms L = Y_(n)ᵀ U_n with Y = A x_{m != n} U_mᵀ
// (ttmc_mTR, template/function/sp_tensor.tpp:236-286), then U_n <- orth(Y_(n) L)
// (ttmc_mTL :288-365 and orth2, operator_double.cpp:218-277): one subspace
// step U_n <- orth(S U_n) with S = Y_(n) Y_(n)ᵀ never formed.  The final core
// is L of the last mode (:110) and the residual sqrt(1 - ||G||^2/||A||^2).
//
// Layout on the device (dt_sp): the COO as given (int32 indices per mode,
// FP32 values) and, per mode n, a copy sorted stably by index n (the
// reference's SpTensor constructor counting sort, sp_tensor.hpp:117-147,
// which it applies to its slice mode): the permutation, the sorted mode-n
// index of every nonzero, the other modes' indices and the values in that
// order, row pointers (I_n + 1), and a compact numbering ("slot") of the
// non-empty rows.  Sorting uses the stable LSD radix sort of dt_sort.hip, so
// the permutation is exactly the reference's.
//
// TTMc (sp_ttmc_dev): Y_c[slot(i), :] = sum over the nonzeros of row i of
// val * (U_a[i_a, :] ⊗ U_b[i_b, :]) for the two other modes a < b (entry
// r_a + R_a r_b, the reference's add_outer_product_all order).  Every wave
// takes one contiguous range of the mode-sorted stream; four nonzeros at a
// time are one v_mfma_f32_16x16x4_f32 (A operand = val * U_a rows, B operand
// = U_b rows, so D accumulates the 16 x 16 outer-product sum of a row
// segment).  A row wholly inside a range is written directly; a row cut by
// range boundaries leaves FP32 partials in carry slots that sp_fixup and
// sp_fixup_chunks add in a fixed order (deterministic).  L = Y_cᵀ U_n is a second, dense pass over Y_c
// (sp_ytu) with per-workgroup partials summed in a fixed order in FP64; the
// factor update (sp_tl_gram, sp_rmul) never waits for the host.
#include "dt_sparse.hpp"
#include "dt_philox.hpp"
#include "dt_chol16.hpp"


#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int SP_E = 256;    // Kronecker row length handled per wave (16 x 16)

int bits_for(int64_t dim) {
    int b = 0;
    while (b < 62 && ((int64_t)1 << b) < dim) b++;
    return b < 1 ? 1 : b;
}

inline dim3 sp_grid(int64_t n) { return dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65535)); }

// ---- generator -------------------------------------------------------------

// key(d) = sum_n idx_n(d) << shift_n, idx_n = first k with cdf_n[k] >= u,
// u = uniform(seed, 7000 + n, d); value(d) = uniform(seed, 7100, d) in FP32.
__global__ void zipf_draw(int64_t ndraws, int ndim, const double *__restrict__ cdf, const int64_t *__restrict__ coff,
                          const int64_t *__restrict__ dims, const int *__restrict__ shift, uint64_t seed,
                          uint64_t *__restrict__ key, int32_t *__restrict__ pos, float *__restrict__ vdraw) {
    for (int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; d < ndraws; d += (int64_t)gridDim.x * blockDim.x) {
        uint64_t k = 0;
        for (int n = 0; n < ndim; n++) {
            const double u = uniform_at(seed, 7000 + (uint64_t)n, (uint64_t)d);
            const double *c = cdf + coff[n];
            int64_t lo = 0, hi = dims[n] - 1;  // cdf[dims-1] == 1 >= u
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (c[mid] >= u)
                    hi = mid;
                else
                    lo = mid + 1;
            }
            k |= (uint64_t)lo << shift[n];
        }
        key[d] = k;
        pos[d] = (int32_t)d;
        vdraw[d] = (float)uniform_at(seed, 7100, (uint64_t)d);
    }
}

__global__ void run_heads(const uint64_t *__restrict__ key, int64_t n, int32_t *__restrict__ flag) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        flag[p] = (p == 0 || key[p] != key[p - 1]) ? 1 : 0;
}

__global__ void scatter_heads(const int32_t *__restrict__ flag, const int32_t *__restrict__ uidx, int64_t n,
                              int32_t *__restrict__ heads) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        if (flag[p]) heads[uidx[p]] = (int32_t)p;
}

constexpr int LONG_RUN = 4096;

// unique entry j: indices from the key, value = FP64 sum of its draws in draw
// order, rounded once to FP32; runs longer than LONG_RUN are queued.
__global__ void merge_runs(const uint64_t *__restrict__ key, const int32_t *__restrict__ pos,
                           const float *__restrict__ vdraw, const int32_t *__restrict__ heads, int64_t nu, int64_t n,
                           int ndim, const int *__restrict__ shift, const int *__restrict__ nbits,
                           int32_t *const *__restrict__ idx, float *__restrict__ val, int32_t *__restrict__ longq,
                           int32_t *__restrict__ nlong) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nu; j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t h = heads[j], e = j + 1 < nu ? heads[j + 1] : n;
        const uint64_t k = key[h];
        for (int m = 0; m < ndim; m++) idx[m][j] = (int32_t)((k >> shift[m]) & ((1ull << nbits[m]) - 1));
        if (e - h > LONG_RUN) {
            longq[atomicAdd(nlong, 1)] = (int32_t)j;
            continue;
        }
        double s = 0.0;
        for (int64_t q = h; q < e; q++) s += (double)vdraw[pos[q]];
        val[j] = (float)s;
    }
}

// one workgroup per long run: 256 strided FP64 partial sums, fixed tree
__global__ __launch_bounds__(256) void merge_long(const int32_t *__restrict__ pos, const float *__restrict__ vdraw,
                                                  const int32_t *__restrict__ heads, int64_t nu, int64_t n,
                                                  const int32_t *__restrict__ longq, float *__restrict__ val) {
    __shared__ double red[256];
    const int64_t j = longq[blockIdx.x];
    const int64_t h = heads[j], e = j + 1 < nu ? heads[j + 1] : n;
    double s = 0.0;
    for (int64_t q = h + threadIdx.x; q < e; q += 256) s += (double)vdraw[pos[q]];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) val[j] = (float)red[0];
}

// ---- mode sort ---------------------------------------------------------------

// flag |= 1 when an index lies outside [0, dim)
__global__ void index_check(const int32_t *__restrict__ idx, int64_t n, int64_t dim, int32_t *__restrict__ flag) {
    int bad = 0;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        bad |= idx[p] < 0 || (int64_t)idx[p] >= dim;
    if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

__global__ void iota32(int32_t *__restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int32_t)i;
}

// rowptr[i] = first position whose sorted key is >= i, rowptr[dim] = n: the
// boundary at position p (keys prev < next around it) writes the pointers of
// rows prev + 1 .. next, so every entry is written once, without atomics.
__global__ void row_bounds(const int32_t *__restrict__ sorted, int64_t n, int64_t dim, int32_t *__restrict__ rowptr) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p <= n; p += (int64_t)gridDim.x * blockDim.x) {
        // clamped to [-1, dim] so that keys out of range cannot write past rowptr
        const int64_t prev = p == 0 ? -1 : min((int64_t)sorted[p - 1], dim);
        const int64_t next = p == n ? dim : min(max((int64_t)sorted[p], (int64_t)-1), dim);
        for (int64_t i = prev + 1; i <= next; i++) rowptr[i] = (int32_t)p;
    }
}

__global__ void nonempty_flags(const int32_t *__restrict__ rowptr, int64_t I, int32_t *__restrict__ f) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < I; i += (int64_t)gridDim.x * blockDim.x)
        f[i] = rowptr[i + 1] > rowptr[i] ? 1 : 0;
}

__global__ void rows_list(const int32_t *__restrict__ f, const int32_t *__restrict__ slot, int64_t I,
                          int32_t *__restrict__ rows) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < I; i += (int64_t)gridDim.x * blockDim.x)
        if (f[i]) rows[slot[i]] = (int32_t)i;
}

template <typename T>
__global__ void gather_by(const T *__restrict__ src, const int32_t *__restrict__ perm, int64_t n, T *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[perm[i]];
}

// ---- TTMc --------------------------------------------------------------------

typedef dt_intx4 intx4;
typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef int intx2 __attribute__((ext_vector_type(2)));
typedef double double2x __attribute__((ext_vector_type(2)));

struct SpModeArgs {
    const int32_t *row;     // sorted mode-n index of each nonzero
    const intx4 *rec;       // per sorted nonzero: {index in mode a, index in mode b, compact row, value bits}
    const int32_t *rowptr;  // I_n + 1
    const int32_t *slot;    // compact row number
    int64_t nnz;
};

__global__ void build_rec(const int32_t *__restrict__ row, const int32_t *__restrict__ ia,
                          const int32_t *__restrict__ ib, const float *__restrict__ val,
                          const int32_t *__restrict__ slot, int64_t n, intx4 *__restrict__ rec) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        rec[p] = intx4{ia[p], ib[p], slot[row[p]], __float_as_int(val[p])};
}

// Sort key of a record: (compact row, index in mode a); the position rides along.
__global__ void rec_keys(const intx4 *__restrict__ rec, int64_t n, int ba, uint64_t *__restrict__ key,
                         int32_t *__restrict__ pos) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const intx4 r = rec[p];
        key[p] = ((uint64_t)(uint32_t)r.z << ba) | (uint32_t)r.x;
        pos[p] = (int32_t)p;
    }
}

__global__ void rec_gather(const intx4 *__restrict__ src, const int32_t *__restrict__ pos, int64_t n,
                           intx4 *__restrict__ dst) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        dst[p] = src[pos[p]];
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ranges per chunk of a long row's chain of carry slots (sp_fixup)
constexpr int SP_FC = 64;

// Per-wave LDS: the staged batch's U_a and U_b rows [nonzero][column] and a
// ring of three batches' (compact row, value) pairs.
constexpr int SP_RING = 3;
struct alignas(16) SpLds {
    float a[64 * 16];
    float b[64 * 16];
    intx2 sv[SP_RING][64];
};

// Record of nonzero p of the range; past P1 the last record of the range is
// loaded (its indices are valid gather addresses) and the (row, value) pair
// is replaced by (-1, 0) where it is staged.  The load is issued for every
// lane and its result is not touched until it is needed, so no wait for it is
// forced early.
__device__ __forceinline__ intx4 sp_rec_load(const intx4 *__restrict__ rec, int64_t p, int64_t P1) {
    return __builtin_nontemporal_load(rec + (p < P1 ? p : P1 - 1));
}

// Gathers of one batch: instruction s, lane (p4, c4) loads columns 4c4..4c4+3
// of U_a[i_a] and U_b[i_b] of nonzero 16 s + p4 (whole 64-byte rows, four lanes
// each).  OFF32: both factors are below 4 GiB, so the address is the factor's
// (scalar) base plus a 32-bit byte offset, one VALU operation per gather.
template <bool OFF32>
__device__ __forceinline__ void sp_issue(const intx4 &z, const float *__restrict__ Ua, const float *__restrict__ Ub,
                                         int p4, int c4, floatx4 (&ga)[4], floatx4 (&gb)[4]) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int src = (16 * s + p4) * 4;
        const int ia = __builtin_amdgcn_ds_bpermute(src, z.x);
        const int ib = __builtin_amdgcn_ds_bpermute(src, z.y);
        if (OFF32) {
            const uint32_t oa = ((uint32_t)ia << 6) + 16u * (uint32_t)c4;
            const uint32_t ob = ((uint32_t)ib << 6) + 16u * (uint32_t)c4;
            ga[s] = *(const floatx4 *)((const char *)Ua + oa);
            gb[s] = *(const floatx4 *)((const char *)Ub + ob);
        } else {
            ga[s] = *(const floatx4 *)(Ua + (int64_t)ia * 16 + 4 * c4);
            gb[s] = *(const floatx4 *)(Ub + (int64_t)ib * 16 + 4 * c4);
        }
    }
}

// Every wave owns the contiguous range [P0, P1) of the mode-sorted records
// (equal sizes, multiples of 64) and walks it in batches of 64 nonzeros.  For
// batch k it stages the gathered rows into LDS (U_a rows multiplied by their
// value), issues the gathers of batch k + 2 into the registers just freed and
// the record load of batch k + 4, then runs batch k from LDS; two register
// sets alternate (the loop body is unrolled by two, so no register holding a
// load in flight is ever copied).  MFMA step t, lane (c, g) takes
// val * U_a[i_a, c] and U_b[i_b, c] of nonzero 4t + g.  A batch inside the
// current row runs its 16 steps with no test between them; a batch where rows
// start runs, row by row, the steps that hold the row's nonzeros with the
// other lanes zeroed and flushes the row before the next one starts.  A
// finished row is written to Y_c when it lies inside the range; otherwise it
// goes to carry slot 0 (row started before P0) or 1 (row continues past P1),
// and sp_fixup adds them.  Four waves per SIMD (128 registers; LDS allows
// four workgroups per CU).
template <bool OFF32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void sp_ttmc16(SpModeArgs a, const float *__restrict__ Ua,
                                                 const float *__restrict__ Ub, float *__restrict__ Yc,
                                                 float *__restrict__ carry, int64_t per_wave, int64_t nwaves,
                                                 int *__restrict__ fix_list, int *__restrict__ fix_count) {
    __shared__ SpLds lds_all[4];
    // the wave index is uniform: readfirstlane lets the compiler keep the
    // range bounds and the batch loop scalar
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    SpLds &L = lds_all[wv];
    const int c = lane & 15, g = lane >> 4;     // MFMA operand layout
    const int p4 = lane >> 2, c4 = lane & 3;    // gather layout
    const int64_t w = (int64_t)blockIdx.x * 4 + wv;
    if (w >= nwaves) return;
    const int64_t P0 = w * per_wave;
    const int64_t P1 = P0 + per_wave < a.nnz ? P0 + per_wave : a.nnz;
    if (P0 >= P1) return;
    const intx4 *__restrict__ rec = a.rec;
    const int first_row = rec[P0].z;
    const bool left_inc = P0 > 0 && rec[P0 - 1].z == first_row;
    const int after = P1 < a.nnz ? rec[P1].z : -1;
    int cur = first_row;
    // partial sums of the current row: two MFMA accumulators in the D layout
    // (lane (c, g), v -> entry 4g + v + 16c)
    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    const int e0 = 4 * g + 16 * c;

    auto flush = [&](bool last) {
        const bool linc = cur == first_row && left_inc;
        const bool rinc = last && after == cur;
        float *dst = (!linc && !rinc) ? Yc + (int64_t)cur * SP_E : carry + (w * 2 + (linc ? 0 : 1)) * SP_E;
        __builtin_nontemporal_store(acc0 + acc1, (floatx4 *)(dst + e0));
    };

    // heads: the lanes where a row starts (bit k: nonzero k), valid: the
    // lanes that are not padding; both found when the batch's records were
    // issued (row_marks), so a batch inside one row reads nothing but its
    // operands here
    auto compute = [&](int ks, uint64_t heads, uint64_t valid) {
        // one base address per operand; step t is a constant offset from it
        const float *pa = L.a + g * 16 + c;
        const float *pb = L.b + g * 16 + c;
        if (heads == 0 && valid == ~0ull) {
            // the whole batch continues the current row: 16 MFMA steps with
            // no test between them
#pragma unroll
            for (int t = 0; t < 16; t++) {
                const float av = pa[64 * t];
                const float bv = pb[64 * t];
                if (t & 1)
                    acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc1, 0, 0, 0);
                else
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc0, 0, 0, 0);
            }
            return;
        }
        // row starts and/or a partial last batch: step by step (operands read
        // four steps at a time), the lanes of the current row in each step,
        // the others masked; at a row start inside a step the finished row is
        // flushed and the step goes on with the next row
        const int nvalid = __popcll(valid);
        for (int t4 = 0; t4 < 4; t4++) {
            if (16 * t4 >= nvalid) break;  // uniform
            float a4[4], b4[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                a4[j] = pa[64 * (4 * t4 + j)];
                b4[j] = pb[64 * (4 * t4 + j)];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int t = 4 * t4 + j;
                const int k = 4 * t + g;
                int lo = 4 * t;
                uint32_t hs = (uint32_t)(heads >> (4 * t)) & 0xfu;
                while (hs) {
                    const int h = 4 * t + __builtin_ctz(hs);
                    if (h > lo)
                        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(k >= lo && k < h ? a4[j] : 0.f, b4[j], acc0, 0, 0,
                                                                    0);
                    flush(false);
                    cur = L.sv[ks][h].x;
                    acc0 = floatx4{0.f, 0.f, 0.f, 0.f};
                    acc1 = floatx4{0.f, 0.f, 0.f, 0.f};
                    lo = h;
                    hs &= hs - 1;
                }
                const int hi = 4 * t + 4 < nvalid ? 4 * t + 4 : nvalid;
                if (hi > lo)
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(k >= lo && k < hi ? a4[j] : 0.f, b4[j], acc0, 0, 0, 0);
            }
        }
    };

    // heads / valid of a batch from its records (row -1 past P1): lane k's
    // predecessor row comes from lane k - 1 by a wave shift, lane 0's is the
    // last row of the batch issued before (the first row of the range for
    // the first batch), which is the current row when the batch is computed
    int lastq = first_row;
    auto row_marks = [&](int row, uint64_t &heads, uint64_t &valid) {
        const int prev = __builtin_amdgcn_update_dpp(lastq, row, 0x138, 0xf, 0xf, false);  // wave_shr:1
        valid = __ballot(row >= 0);
        heads = __ballot(row >= 0 && row != prev);
        lastq = __builtin_amdgcn_readlane(row, 63);
    };

    // batch k (first nonzero b0, ring slot ks) with register set (z, ga, gb,
    // hd, vd): z holds the records of batch k + 2, ga/gb the gathers of batch
    // k, hd/vd its heads and valid lanes
    auto step = [&](int64_t b0, int ks, intx4 &z, floatx4 (&ga)[4], floatx4 (&gb)[4], uint64_t &hd,
                    uint64_t &vd) {
        wave_sync();  // the previous batch's LDS reads are done
        // U_a rows are staged already scaled by their nonzero's value
        // (padding has value 0)
        const intx2 *psv = &L.sv[ks][p4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const float v = __int_as_float(psv[16 * s].y);
            *(floatx4 *)(L.a + (16 * s + p4) * 16 + 4 * c4) = ga[s] * v;
            *(floatx4 *)(L.b + (16 * s + p4) * 16 + 4 * c4) = gb[s];
        }
        const int ks2 = ks == 0 ? 2 : ks - 1;  // (ks + 2) mod 3
        const bool ok = b0 + 128 + lane < P1;
        const int rown = ok ? z.z : -1;
        L.sv[ks2][lane] = intx2{rown, ok ? z.w : 0};
        const uint64_t h = hd, v = vd;
        row_marks(rown, hd, vd);
        sp_issue<OFF32>(z, Ua, Ub, p4, c4, ga, gb);
        z = sp_rec_load(rec, b0 + 256 + lane, P1);
        wave_sync();
        compute(ks, h, v);
    };

    intx4 zX = sp_rec_load(rec, P0 + lane, P1);
    intx4 zY = sp_rec_load(rec, P0 + 64 + lane, P1);
    floatx4 gaX[4], gbX[4], gaY[4], gbY[4];
    uint64_t hX, vX, hY, vY;
    {
        const bool ok = P0 + lane < P1;
        const int row = ok ? zX.z : -1;
        L.sv[0][lane] = intx2{row, ok ? zX.w : 0};
        row_marks(row, hX, vX);
    }
    sp_issue<OFF32>(zX, Ua, Ub, p4, c4, gaX, gbX);
    zX = sp_rec_load(rec, P0 + 128 + lane, P1);
    {
        const bool ok = P0 + 64 + lane < P1;
        const int row = ok ? zY.z : -1;
        L.sv[1][lane] = intx2{row, ok ? zY.w : 0};
        row_marks(row, hY, vY);
    }
    sp_issue<OFF32>(zY, Ua, Ub, p4, c4, gaY, gbY);
    zY = sp_rec_load(rec, P0 + 192 + lane, P1);
    // a counted loop over pairs of batches (padding past P1 has row -1 and
    // value 0, so the last batch needs no bounds test), then the odd batch
    const int64_t nb = (P1 - P0 + 63) / 64;
    int ks = 0;
    int64_t b0 = P0;
    for (int64_t i = 0; i < nb / 2; i++) {
        step(b0, ks, zX, gaX, gbX, hX, vX);
        ks = ks == 2 ? 0 : ks + 1;
        step(b0 + 64, ks, zY, gaY, gbY, hY, vY);
        ks = ks == 2 ? 0 : ks + 1;
        b0 += 128;
    }
    if (nb & 1) step(b0, ks, zX, gaX, gbX, hX, vX);
    flush(true);
    // ranges with work for sp_fixup: the owner of a row that continues past
    // P1, and the start of every SP_FC-th range inside a longer row's chain
    const bool owner = after == cur && !(cur == first_row && left_inc);
    bool chunk = false;
    if (left_inc) {
        const int64_t o = a.rowptr[a.row[P0]] / per_wave;
        chunk = (w - o) % SP_FC == 0;
    }
    if ((owner || chunk) && lane == 0) fix_list[atomicAdd(fix_count, 1)] = (int)w;
}

// Rows cut by wave ranges: the row's partials are slot 1 of the range that
// holds its first nonzero (the owner) and slot 0 of every later range it
// reaches.  A Zipf-hot row spans thousands of ranges, so its chain is cut into
// chunks of SP_FC ranges summed in parallel (sp_fixup) and the chunk sums are
// added by the owner (sp_fixup_chunks); chains of one chunk are finished in
// the first kernel.  Every sum runs in a fixed order for a given range size.
// Range w's chunk of the chain of row r that starts at range s0 = w (owner
// when w holds the row's first nonzero, else a chunk start inside the chain):
// the sum of slots [w, min(w + SP_FC, we + 1)) by one wave (lane t: entries
// 4t .. 4t + 3), sixteen slots in flight (a hot row's chunk takes four
// rounds of loads), added in a fixed order.
__device__ __forceinline__ floatx4 sp_sum16(const floatx4 (&acc)[16]) {
    floatx4 h[8];
#pragma unroll
    for (int u = 0; u < 8; u++) h[u] = acc[2 * u] + acc[2 * u + 1];
    return ((h[0] + h[1]) + (h[2] + h[3])) + ((h[4] + h[5]) + (h[6] + h[7]));
}

__device__ __forceinline__ floatx4 sp_chunk_sum(const float *__restrict__ carry, int64_t w, bool owner, int64_t we) {
    const int t = threadIdx.x;
    const int64_t qe = w + SP_FC - 1 < we ? w + SP_FC - 1 : we;
    auto slot = [&](int64_t x) {
        return *(const floatx4 *)(carry + (x * 2 + (owner && x == w ? 1 : 0)) * SP_E + 4 * t);
    };
    floatx4 acc[16];
#pragma unroll
    for (int u = 0; u < 16; u++) acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
    int64_t q = w;
    for (; q + 15 <= qe; q += 16)
#pragma unroll
        for (int u = 0; u < 16; u++) acc[u] += slot(q + u);
#pragma unroll
    for (int u = 0; u < 15; u++)
        if (q + u <= qe) acc[u] += slot(q + u);
    return sp_sum16(acc);
}

// One wave per entry of the fix-up list (most chains are two or three slots
// long, so the entries are many and short).
__global__ __launch_bounds__(64) void sp_fixup(SpModeArgs a, const float *__restrict__ carry, int64_t per_wave,
                                               float *__restrict__ csum, float *__restrict__ Yc,
                                               const int *__restrict__ fix_list, const int *__restrict__ fix_count) {
    const int nlist = *fix_count;
    const int t = threadIdx.x;
    for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
        const int64_t w = fix_list[li];
        const int64_t P0 = w * per_wave;
        const int64_t P1 = P0 + per_wave < a.nnz ? P0 + per_wave : a.nnz;
        // role 1: owner of the chain of its last row
        {
            const int r = a.row[P1 - 1];
            const int64_t rs = a.rowptr[r], re = a.rowptr[r + 1];
            if (re > P1 && rs >= P0) {  // uniform
                const int64_t we = (re - 1) / per_wave;
                const floatx4 v = sp_chunk_sum(carry, w, true, we);
                float *dst = we - w < SP_FC ? Yc + (int64_t)a.slot[r] * SP_E : csum + (w * 2 + 1) * SP_E;
                *(floatx4 *)(dst + 4 * t) = v;
            }
        }
        // role 2: start of a later chunk of the chain of its first row
        {
            const int r = a.row[P0];
            const int64_t rs = a.rowptr[r], re = a.rowptr[r + 1];
            if (rs < P0) {  // uniform
                const int64_t o = rs / per_wave;
                if ((w - o) % SP_FC == 0) {
                    const int64_t we = (re - 1) / per_wave;
                    *(floatx4 *)(csum + (w * 2) * SP_E + 4 * t) = sp_chunk_sum(carry, w, false, we);
                }
            }
        }
    }
}

// Owners of chains longer than one chunk: the chunk sums in chain order.
__global__ __launch_bounds__(64) void sp_fixup_chunks(SpModeArgs a, const float *__restrict__ csum,
                                                      int64_t per_wave, float *__restrict__ Yc,
                                                      const int *__restrict__ fix_list,
                                                      const int *__restrict__ fix_count) {
    const int nlist = *fix_count;
    for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
        const int64_t w = fix_list[li];
        const int64_t P0 = w * per_wave;
        const int64_t P1 = P0 + per_wave < a.nnz ? P0 + per_wave : a.nnz;
        const int r = a.row[P1 - 1];
        const int64_t rs = a.rowptr[r], re = a.rowptr[r + 1];
        if (re <= P1 || rs < P0) continue;  // not an owner
        const int64_t we = (re - 1) / per_wave;
        if (we - w < SP_FC) continue;  // finished by sp_fixup
        const int t = threadIdx.x;
        auto chunk = [&](int64_t s) { return *(const floatx4 *)(csum + (s * 2) * SP_E + 4 * t); };
        // sixteen chunk sums in flight, added in a fixed order
        floatx4 acc[16];
#pragma unroll
        for (int u = 0; u < 16; u++) acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
        int64_t s = w + SP_FC;
        for (; s + 15 * SP_FC <= we; s += 16 * SP_FC)
#pragma unroll
            for (int u = 0; u < 16; u++) acc[u] += chunk(s + u * SP_FC);
#pragma unroll
        for (int u = 0; u < 15; u++)
            if (s + u * SP_FC <= we) acc[u] += chunk(s + u * SP_FC);
        const floatx4 tot = *(const floatx4 *)(csum + (w * 2 + 1) * SP_E + 4 * t) + sp_sum16(acc);
        *(floatx4 *)(Yc + (int64_t)a.slot[r] * SP_E + 4 * t) = tot;
    }
}

// Partial L = Y_cᵀ U_n[rows] (256 x 16) per 256-thread workgroup on FP32
// MFMA.  Each wave walks its own contiguous run [s0, s1) of compact rows four
// at a time (the MFMA's K): lane (c, g) loads entries 64q + 4c .. 64q + 4c + 3
// (q = 0..3) of row s + g, so each load instruction reads four 256-byte
// contiguous pieces, and U_n[rows[s + g], c].  MFMA t = 4q + r takes
// A[i][g] = Y_c[s + g, 64q + 4i + r] and B[g][j] = U_n[rows[s + g], j], so
// accumulator t ends as L[64q + 4i + r, j] with i = 4g + v (register v) and
// j = c.  Two register sets alternate over groups of four rows (one group's
// loads are in flight while the other's MFMAs run) and each set's row indices
// are loaded one of its groups ahead.  Rows past s1 are read as row s1 - 1
// with their U_n value zeroed.  The four waves' sums are then added in wave
// order through LDS.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void sp_ytu(
    const float *__restrict__ Yc, const int32_t *__restrict__ rows, int64_t nrows, const float *__restrict__ Un,
    int64_t rows_per_wave, float *__restrict__ Lwg) {
    __shared__ floatx4 red[16][64];  // [t][lane]: the running sum of the waves
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = lane & 15, g = lane >> 4;
    const int64_t s0 = ((int64_t)blockIdx.x * 4 + wv) * rows_per_wave;
    const int64_t s1 = s0 + rows_per_wave < nrows ? s0 + rows_per_wave : nrows;
    floatx4 acc[16];
#pragma unroll
    for (int t = 0; t < 16; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (s0 < s1) {  // uniform
        const int64_t last = s1 - 1;
        auto row_of = [&](int64_t s) { return s + g <= last ? s + g : last; };
        auto ld_y = [&](int64_t s, floatx4(&y)[4]) {
            const float *p = Yc + row_of(s) * SP_E + 4 * c;
#pragma unroll
            for (int q = 0; q < 4; q++) y[q] = __builtin_nontemporal_load((const floatx4 *)(p + 64 * q));
        };
        auto ld_u = [&](int64_t s, int r) {
            const float u = Un[(int64_t)r * 16 + c];
            return s + g <= last ? u : 0.f;
        };
        auto compute = [&](const floatx4(&y)[4], float u) {
#pragma unroll
            for (int t = 0; t < 16; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(y[t >> 2][t & 3], u, acc[t], 0, 0, 0);
        };
        const int64_t ngrp = (s1 - s0 + 3) / 4;
        floatx4 yX[4], yY[4];
        const int rX0 = rows[row_of(s0)], rY0 = rows[row_of(s0 + 4)];
        ld_y(s0, yX);
        float uX = ld_u(s0, rX0);
        ld_y(s0 + 4, yY);
        float uY = ld_u(s0 + 4, rY0);
        int rX = rows[row_of(s0 + 8)], rY = rows[row_of(s0 + 12)];
        int64_t s = s0;
        for (int64_t k = 0; k < ngrp / 2; k++) {
            compute(yX, uX);
            ld_y(s + 8, yX);
            uX = ld_u(s + 8, rX);
            rX = rows[row_of(s + 16)];
            compute(yY, uY);
            ld_y(s + 12, yY);
            uY = ld_u(s + 12, rY);
            rY = rows[row_of(s + 20)];
            s += 8;
        }
        if (ngrp & 1) compute(yX, uX);
    }
    // waves 0..3 in order: each adds its sums to the running sum in LDS
    for (int w = 0; w < 4; w++) {
        if (wv == w) {
#pragma unroll
            for (int t = 0; t < 16; t++) red[t][lane] = w > 0 ? red[t][lane] + acc[t] : acc[t];
        }
        __syncthreads();
    }
    // partial, row-major 256 x 16: entry e = 64q + 4i + r (accumulator 4q + r), column j
    float *L = Lwg + (int64_t)blockIdx.x * SP_E * 16;
    for (int x = threadIdx.x; x < SP_E * 16; x += 256) {
        const int j = x & 15, e = x >> 4, i = (e >> 2) & 15, t = 4 * (e >> 6) + (e & 3);
        L[x] = red[t][j + 16 * (i >> 2)][i & 3];
    }
}

// L (256 x 16, FP64) = sum over workgroups of the partials, fixed order: 16
// entries per workgroup, 16 groups of threads each summing every 16th partial,
// then the groups added in order.
__global__ __launch_bounds__(256) void sp_lreduce(const float *__restrict__ Lwg, int nwg, double *__restrict__ L) {
    __shared__ double red[16][17];
    const int c = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int i = blockIdx.x * 16 + c;
    // eight independent sums per thread so that eight loads are in flight
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int w = grp;
    for (; w + 7 * 16 < nwg; w += 8 * 16)
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] += (double)Lwg[(int64_t)(w + 16 * u) * SP_E * 16 + i];
    for (; w < nwg; w += 16) a[0] += (double)Lwg[(int64_t)w * SP_E * 16 + i];
    red[grp][c] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    __syncthreads();
    if (grp == 0) {
        double t = 0.0;
        for (int g = 0; g < 16; g++) t += red[g][c];
        L[i] = t;
    }
}

// Y (I x 256 row-major FP32, zeros for empty rows) from Y_c.
__global__ void sp_expand(const float *__restrict__ Yc, const int32_t *__restrict__ rows, int64_t nrows,
                          float *__restrict__ Y) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nrows * SP_E; e += (int64_t)gridDim.x * blockDim.x)
        Y[(int64_t)rows[e / SP_E] * SP_E + e % SP_E] = Yc[e];
}

// ---- factor update without host round trips ----------------------------------
// U_n <- orth(Y_(n) L) as shifted Cholesky-QR3 (Fukaya et al. 2020): one shifted
// pass, then two plain ones.  Every pass streams U once on FP64 MFMA and leaves
// the Gram of its output as per-workgroup partials (fixed order), so no step
// waits for the host; a Cholesky breakdown sets a device flag that the sweep
// reads with its residual.

// Partial Gram of a wave's 16 x 16 FP64 MFMA accumulator (lane (c, g), register
// v = G[4v + g][c]): the four waves are added in wave order and entry (a, b),
// a, b < R, goes to part[blockIdx.x * R * R + a + R b].
__device__ __forceinline__ void sp_gram_out(doublex4 gacc, int R, double *__restrict__ part) {
    __shared__ double red[4][4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < 4; v++) red[wv][v][lane] = gacc[v];
    __syncthreads();
    const int v = threadIdx.x >> 6, c = lane & 15, g = lane >> 4;
    const double sum = ((red[0][v][lane] + red[1][v][lane]) + red[2][v][lane]) + red[3][v][lane];
    const int a = 4 * v + g;
    if (a < R && c < R) part[(int64_t)blockIdx.x * R * R + a + R * c] = sum;
}

// U (I x R, FP64 column-major) = Y_(n) L for every row i of the mode (rows
// without nonzeros give 0) and the partial Grams of U.  A wave takes runs of
// 16-row tiles: lane (i, g) reads row 16t + i of Y (its compact copy) as
// float4s 4q + g, one FP32 MFMA step per float4 component against L (FP32 in
// LDS) leaves W[16t + 4g + v][i] in register v; those values, widened to FP64,
// are fed to the Gram's FP64 MFMA as both operands and stored column by
// column through a per-wave LDS tile.
__global__ __launch_bounds__(256) void sp_tl_gram(const float *__restrict__ Yc, const int32_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ slot, int64_t I, int R,
                                                  const double *__restrict__ L, int64_t tiles_per_wave,
                                                  double *__restrict__ U, double *__restrict__ part) {
    __shared__ float Ls[SP_E][20];
    __shared__ double Wt[4][16][18];  // per wave: the tile by column, rows contiguous
    for (int e = threadIdx.x; e < SP_E * 16; e += 256) Ls[e / 16][e % 16] = (float)L[e];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lane & 15, g = lane >> 4;
    const int64_t ntiles = (I + 15) / 16;
    const int64_t t0 = ((int64_t)blockIdx.x * 4 + wv) * tiles_per_wave;
    const int64_t t1 = t0 + tiles_per_wave < ntiles ? t0 + tiles_per_wave : ntiles;
    doublex4 gacc = {0.0, 0.0, 0.0, 0.0};
    for (int64_t t = t0; t < t1; t++) {
        const int64_t r = t * 16 + i;
        const bool ok = r < I && rowptr[r + 1] > rowptr[r];
        const floatx4 *yr = (const floatx4 *)(Yc + (int64_t)(ok ? slot[r] : 0) * SP_E);
        floatx4 y[16];
#pragma unroll
        for (int q = 0; q < 16; q++) y[q] = ok ? yr[4 * q + g] : floatx4{0.f, 0.f, 0.f, 0.f};
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 16; q++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(y[q][j], Ls[16 * q + 4 * g + j][i], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double w = (double)acc[v];
            Wt[wv][i][4 * g + v] = w;
            gacc = __builtin_amdgcn_mfma_f64_16x16x4f64(w, w, gacc, 0, 0, 0);
        }
        wave_sync();
        {
            // the tile goes out by column: lane l stores rows 4 (l % 4) .. +3 of
            // column l / 4, so four lanes write one 128-byte column piece
            // (16-byte stores when the column start is 16-byte aligned)
            const int col = lane >> 2, r4 = 4 * (lane & 3);
            const int64_t r0 = t * 16 + r4;
            const double *src = &Wt[wv][col][r4];
            double *dst = U + r0 + I * col;
            if (col < R) {
                if (r0 + 3 < I && (I & 1) == 0) {
                    *(double2x *)dst = double2x{src[0], src[1]};
                    *(double2x *)(dst + 2) = double2x{src[2], src[3]};
                } else {
                    for (int q = 0; q < 4; q++)
                        if (r0 + q < I) dst[q] = src[q];
                }
            }
        }
        wave_sync();  // the tile buffer is rewritten by the next tile
    }
    sp_gram_out(gacc, R, part);
}

// One Cholesky-QR pass: U <- U R^{-1} in place (U: I x R FP64 column-major,
// R <= 16) with R = chol(G + shift), G the summed Gram of U (R x R), the
// shift 11 (I R + R(R+1)) u tr(G) when `shifted`.  Every workgroup factors G
// itself (wave 0, chol16_rinv: a few microseconds, no launch of its own) and
// all of them get the same R^{-1}; a non-positive pivot sets bit 4 of flag.
// The product runs on FP64 MFMA: for a 16-row tile lane (c, g) feeds
// U[r0 + c][4s + g] as A and R^{-1}[4s + g][c] as B of step s and receives
// Unew[r0 + 4v + g][c] in register v.  GRAM: partial Grams of the new U.
// VIEW: also the FP32 row-major copy, 16 columns (zero past R).
template <bool GRAM, bool VIEW>
__global__ __launch_bounds__(256) void sp_rmul(double *__restrict__ U, int64_t I, int R,
                                               const double *__restrict__ Gs, int shifted, int *__restrict__ flag,
                                               int64_t tiles_per_wave, double *__restrict__ part,
                                               float *__restrict__ Ur) {
    __shared__ double Wt[4][16][18];  // per wave: the new tile by column
    __shared__ double Gl[16][17], Ril[16][17];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    {
        const int i = threadIdx.x >> 4, k = threadIdx.x & 15;
        Gl[i][k] = (i < R && k < R) ? Gs[i + R * k] : 0.0;
    }
    __syncthreads();
    chol16_rinv(Gl, Ril, R, (int)I, shifted, flag);  // wave 0; the others return at once
    __syncthreads();
    double b[4];
#pragma unroll
    for (int s = 0; s < 4; s++) b[s] = (4 * s + g < R && c < R) ? Ril[4 * s + g][c] : 0.0;
    const int64_t ntiles = (I + 15) / 16;
    const int64_t t0 = ((int64_t)blockIdx.x * 4 + wv) * tiles_per_wave;
    const int64_t t1 = t0 + tiles_per_wave < ntiles ? t0 + tiles_per_wave : ntiles;
    doublex4 gacc = {0.0, 0.0, 0.0, 0.0};
    for (int64_t t = t0; t < t1; t++) {
        const int64_t r0 = t * 16;
        double a[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int k = 4 * s + g;
            a[s] = (k < R && r0 + c < I) ? U[(r0 + c) + I * k] : 0.0;
        }
        doublex4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 4; s++) d = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], d, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int64_t rr = r0 + 4 * v + g;
            Wt[wv][c][4 * v + g] = d[v];
            if (VIEW && rr < I) Ur[rr * 16 + c] = (float)d[v];
            if (GRAM) gacc = __builtin_amdgcn_mfma_f64_16x16x4f64(d[v], d[v], gacc, 0, 0, 0);
        }
        wave_sync();
        {
            // column-wise stores as in sp_tl_gram: lane l writes rows
            // 4 (l % 4) .. +3 of column l / 4
            const int col = lane >> 2, r4 = 4 * (lane & 3);
            const int64_t rb = r0 + r4;
            const double *src = &Wt[wv][col][r4];
            double *dst = U + rb + I * col;
            if (col < R) {
                if (rb + 3 < I && (I & 1) == 0) {
                    *(double2x *)dst = double2x{src[0], src[1]};
                    *(double2x *)(dst + 2) = double2x{src[2], src[3]};
                } else {
                    for (int q = 0; q < 4; q++)
                        if (rb + q < I) dst[q] = src[q];
                }
            }
        }
        wave_sync();
    }
    if (GRAM) sp_gram_out(gacc, R, part);
}

// row-major FP32 view with 16 columns, columns R..15 zero; one row per thread
__global__ void rows16_view(const double *__restrict__ U, int64_t I, int R, float *__restrict__ Ur) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < I; i += (int64_t)gridDim.x * blockDim.x) {
        float v[16];
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = r < R ? (float)U[i + I * r] : 0.f;
        floatx4 *dst = (floatx4 *)(Ur + i * 16);
#pragma unroll
        for (int q = 0; q < 4; q++) dst[q] = floatx4{v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
    }
}

__global__ void sumsq_f32_sp(const float *__restrict__ v, int64_t n, double *__restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += (double)v[i] * v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace

// ---- the sparse tensor ---------------------------------------------------------

static void *sp_alloc(dt_sp *sp, size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) return nullptr;
    sp->owned.push_back(p);
    return p;
}

// temp storage of the radix sort and the scans (grow-only workspace slot)
static void *sort_tmp(dt_ctx *ctx, size_t bytes) { return dt_ws(ctx, WS_SPARSE3, std::max<size_t>(bytes, 256)); }

static int sort_mode_internal(dt_ctx *ctx, int64_t nnz, const int32_t *keys, int64_t dim, int32_t *perm,
                              int32_t *sorted_keys, int32_t *rowptr) {
    int32_t *iota = (int32_t *)dt_ws(ctx, WS_SPARSE2, (size_t)nnz * 4 + 256);
    if (!iota) return DT_ENOMEM;
    DT_LAUNCH("sp_sort", iota32, sp_grid(nnz), dim3(256), 0, iota, nnz);
    const int nb = bits_for(dim);
    void *tmp = sort_tmp(ctx, dt_radix_sort_tmp_bytes(nnz, 4));
    if (!tmp) return DT_ENOMEM;
    DT_TRY(dt_radix_sort_u32(ctx, (const uint32_t *)keys, (uint32_t *)sorted_keys, iota, perm, nnz, nb, tmp));
    // row pointers from the sorted keys (int32: nnz < 2^31)
    DT_LAUNCH("sp_sort", row_bounds, sp_grid(nnz + 1), dim3(256), 0, (const int32_t *)sorted_keys, nnz, dim, rowptr);
    return DT_OK;
}

static int sort_rows_by_a(dt_ctx *ctx, int64_t nnz, int ba, int bs, intx4 *rec) {
    void *buf = nullptr;
    const size_t bytes = (size_t)nnz * (8 + 8 + 4 + 4 + 16) + 1024;
    if (hipMalloc(&buf, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return dt_fail(ctx, DT_ENOMEM, "dt_sp_create: hipMalloc(%zu) for the record sort failed", bytes);
    }
    uint64_t *k1 = (uint64_t *)buf, *k2 = k1 + nnz;
    int32_t *p1 = (int32_t *)(k2 + nnz), *p2 = p1 + nnz;
    intx4 *tmp = (intx4 *)(((uintptr_t)(p2 + nnz) + 255) & ~(uintptr_t)255);
    int st = DT_OK;
    hipLaunchKernelGGL(rec_keys, sp_grid(nnz), dim3(256), 0, ctx->stream, rec, nnz, ba, k1, p1);
    void *ctmp = sort_tmp(ctx, dt_radix_sort_tmp_bytes(nnz, 8));
    if (!ctmp) st = DT_ENOMEM;
    if (st == DT_OK) st = dt_radix_sort_u64(ctx, k1, k2, p1, p2, nnz, ba + bs, ctmp);
    if (st == DT_OK) {
        hipLaunchKernelGGL(rec_gather, sp_grid(nnz), dim3(256), 0, ctx->stream, (const intx4 *)rec, p2, nnz, tmp);
        if (hipMemcpyAsync(rec, tmp, (size_t)nnz * 16, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            st = DT_EHIP;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && st == DT_OK) st = DT_EHIP;
    (void)hipFree(buf);
    if (st != DT_OK && ctx->err.empty()) return dt_fail(ctx, st, "dt_sp_create: record sort failed");
    return st;
}

extern "C" int dt_sp_sort_mode(dt_ctx *ctx, int64_t nnz, const int32_t *keys, size_t dim, int32_t *perm,
                               int32_t *rowptr) {
    DT_CHECK_ARG(ctx && keys && perm && rowptr, "dt_sp_sort_mode: null argument");
    DT_CHECK_ARG(nnz >= 1 && nnz < ((int64_t)1 << 31) && dim >= 1 && dim < ((size_t)1 << 31),
                 "dt_sp_sort_mode: nnz %lld / dim %zu out of range", (long long)nnz, dim);
    int32_t *sk = (int32_t *)dt_ws(ctx, WS_SPARSE0, (size_t)nnz * 4 + 256);
    if (!sk) return DT_ENOMEM;
    return sort_mode_internal(ctx, nnz, keys, (int64_t)dim, perm, sk, rowptr);
}

extern "C" int dt_sp_gen_zipf(dt_ctx *ctx, int ndim, const size_t *dims, int64_t ndraws, double exponent,
                              uint64_t seed, int32_t *const *idx_out, float *val_out, int64_t *nnz_out) {
    DT_CHECK_ARG(ctx && dims && idx_out && val_out && nnz_out, "dt_sp_gen_zipf: null argument");
    DT_CHECK_ARG(ndim >= 1 && ndim <= DT_MAX_NDIM, "dt_sp_gen_zipf: bad ndim");
    DT_CHECK_ARG(ndraws >= 1 && ndraws < ((int64_t)1 << 31), "dt_sp_gen_zipf: draws out of range");
    int shift[DT_MAX_NDIM], nbits[DT_MAX_NDIM], tot = 0;
    int64_t dd[DT_MAX_NDIM], coff[DT_MAX_NDIM], ctot = 0;
    for (int n = 0; n < ndim; n++) {
        DT_CHECK_ARG(dims[n] >= 1 && dims[n] < ((size_t)1 << 31), "dt_sp_gen_zipf: bad dim %d", n);
        DT_CHECK_ARG(idx_out[n], "dt_sp_gen_zipf: index output %d missing", n);
        dd[n] = (int64_t)dims[n];
        nbits[n] = bits_for(dd[n]);
        shift[n] = tot;
        tot += nbits[n];
        coff[n] = ctot;
        ctot += dd[n];
    }
    DT_CHECK_ARG(tot <= 64, "dt_sp_gen_zipf: composite key needs %d > 64 bits", tot);
    // CDF tables on the host: P(k) ~ (k+1)^-s, sequential FP64 sums
    std::vector<double> cdf((size_t)ctot);
    for (int n = 0; n < ndim; n++) {
        double s = 0.0;
        for (int64_t k = 0; k < dd[n]; k++) {
            s += std::pow((double)(k + 1), -exponent);
            cdf[(size_t)(coff[n] + k)] = s;
        }
        for (int64_t k = 0; k < dd[n]; k++) cdf[(size_t)(coff[n] + k)] /= s;
    }
    double *dcdf = (double *)dt_ws(ctx, WS_SPARSE0, (size_t)ctot * 8 + 256);
    int64_t *dmeta = (int64_t *)dt_ws(ctx, WS_SPARSE1, 64 * 8 * 4);
    if (!dcdf || !dmeta) return DT_ENOMEM;
    int64_t *d_coff = dmeta, *d_dims = dmeta + DT_MAX_NDIM;
    int *d_shift = (int *)(dmeta + 2 * DT_MAX_NDIM), *d_nbits = d_shift + DT_MAX_NDIM;
    int32_t **d_idx = (int32_t **)(dmeta + 4 * DT_MAX_NDIM);
    int32_t *d_nlong = (int32_t *)(dmeta + 5 * DT_MAX_NDIM);
    DT_HIP(hipMemcpyAsync(dcdf, cdf.data(), (size_t)ctot * 8, hipMemcpyHostToDevice, ctx->stream));
    DT_HIP(hipMemcpyAsync(d_coff, coff, sizeof(coff), hipMemcpyHostToDevice, ctx->stream));
    DT_HIP(hipMemcpyAsync(d_dims, dd, sizeof(dd), hipMemcpyHostToDevice, ctx->stream));
    DT_HIP(hipMemcpyAsync(d_shift, shift, sizeof(shift), hipMemcpyHostToDevice, ctx->stream));
    DT_HIP(hipMemcpyAsync(d_nbits, nbits, sizeof(nbits), hipMemcpyHostToDevice, ctx->stream));
    DT_HIP(hipMemcpyAsync(d_idx, idx_out, sizeof(int32_t *) * ndim, hipMemcpyHostToDevice, ctx->stream));
    DT_TRY(dt_zero(ctx, d_nlong, 4));
    // draws: keys, positions, values; then a stable radix sort of (key, position)
    const int64_t n = ndraws;
    char *buf = (char *)dt_ws(ctx, WS_GEN, (size_t)n * (8 + 8 + 4 + 4 + 4 + 4 + 4 + 4) + 4096);
    if (!buf) return DT_ENOMEM;
    uint64_t *key = (uint64_t *)buf;
    uint64_t *key2 = key + n;
    int32_t *pos = (int32_t *)(key2 + n);
    int32_t *pos2 = pos + n;
    float *vdraw = (float *)(pos2 + n);
    int32_t *flag = (int32_t *)(vdraw + n);
    int32_t *uidx = flag + n;
    int32_t *heads = uidx + n;
    DT_LAUNCH("sp_gen", zipf_draw, sp_grid(n), dim3(256), 0, n, ndim, dcdf, d_coff, d_dims, d_shift, seed, key, pos,
              vdraw);
    void *tmp = sort_tmp(ctx, std::max(dt_radix_sort_tmp_bytes(n, 8), dt_scan_tmp_bytes(n)));
    if (!tmp) return DT_ENOMEM;
    DT_TRY(dt_radix_sort_u64(ctx, key, key2, pos, pos2, n, tot, tmp));
    DT_LAUNCH("sp_gen", run_heads, sp_grid(n), dim3(256), 0, key2, n, flag);
    DT_TRY(dt_exclusive_scan_i32(ctx, flag, uidx, n, (int32_t *)tmp));
    int32_t last[2];
    DT_HIP(hipMemcpyAsync(&last[0], uidx + n - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    DT_HIP(hipMemcpyAsync(&last[1], flag + n - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t nu = (int64_t)last[0] + last[1];
    DT_LAUNCH("sp_gen", scatter_heads, sp_grid(n), dim3(256), 0, flag, uidx, n, heads);
    int32_t *longq = uidx;  // reused: the run numbering is no longer needed
    DT_LAUNCH("sp_gen", merge_runs, sp_grid(nu), dim3(256), 0, key2, pos2, vdraw, heads, nu, n, ndim, d_shift,
              d_nbits, d_idx, val_out, longq, d_nlong);
    int32_t nlong = 0;
    DT_HIP(hipMemcpyAsync(&nlong, d_nlong, 4, hipMemcpyDeviceToHost, ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    if (nlong > 0)
        DT_LAUNCH("sp_gen", merge_long, dim3((unsigned)nlong), dim3(256), 0, pos2, vdraw, heads, nu, n, longq, val_out);
    *nnz_out = nu;
    return DT_OK;
}

extern "C" int dt_sp_destroy(dt_sp *sp) {
    if (!sp) return DT_OK;
    for (void *p : sp->owned) (void)hipFree(p);
    delete sp;
    return DT_OK;
}

static int sp_create_impl(dt_ctx *ctx, int ndim, const size_t *dims, int64_t nnz, const int32_t *const *idx,
                          const float *val, const double *val64, dt_sp **out) {
    DT_CHECK_ARG(ctx && dims && idx && (val || val64) && out, "dt_sp_create: null argument");
    DT_CHECK_ARG(ndim >= 2 && ndim <= DT_MAX_NDIM, "dt_sp_create: bad ndim %d", ndim);
    DT_CHECK_ARG(nnz >= 1 && nnz < ((int64_t)1 << 31), "dt_sp_create: nnz out of range");
    dt_sp *sp = new dt_sp();
    sp->ndim = ndim;
    sp->nnz = nnz;
    int st = DT_OK;
    // slice mode: the largest mode, first on ties (sp_tensor.hpp:107-115)
    int64_t best = 0;
    for (int n = 0; n < ndim; n++) {
        sp->dims[n] = (int64_t)dims[n];
        if (best < sp->dims[n]) {
            best = sp->dims[n];
            sp->slice_mode = n;
        }
    }
    auto fail = [&](int code, const char *msg) {
        dt_sp_destroy(sp);
        return dt_fail(ctx, code, "%s", msg);
    };
    for (int n = 0; n < ndim; n++) {
        sp->idx[n] = (int32_t *)sp_alloc(sp, (size_t)nnz * 4);
        if (!sp->idx[n]) return fail(DT_ENOMEM, "dt_sp_create: out of memory");
        if (hipMemcpyAsync(sp->idx[n], idx[n], (size_t)nnz * 4, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            return fail(DT_EHIP, "dt_sp_create: copy failed");
    }
    {
        // every index inside its mode (the kernels gather and scatter by them)
        int32_t *flag = (int32_t *)dt_ws(ctx, WS_STATUS, 64);
        if (!flag) return fail(DT_ENOMEM, "dt_sp_create: out of memory");
        if (dt_zero(ctx, flag, 4) != DT_OK) return fail(DT_EHIP, "dt_sp_create: zero failed");
        for (int n = 0; n < ndim; n++)
            hipLaunchKernelGGL(index_check, sp_grid(nnz), dim3(256), 0, ctx->stream, sp->idx[n], nnz, sp->dims[n], flag);
        int32_t hflag = 0;
        if (hipMemcpyAsync(&hflag, flag, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return fail(DT_EHIP, "dt_sp_create: index check failed");
        if (hflag) {
            dt_sp_destroy(sp);
            return dt_fail(ctx, DT_EINVAL, "dt_sp_create: an index lies outside its mode's range");
        }
    }
    sp->val = (float *)sp_alloc(sp, (size_t)nnz * 4);
    if (!sp->val) return fail(DT_ENOMEM, "dt_sp_create: out of memory");
    if (val64) {
        // FP64 values kept for the FP64 path; the FP32 copy is their rounding
        sp->val64 = (double *)sp_alloc(sp, (size_t)nnz * 8);
        if (!sp->val64) return fail(DT_ENOMEM, "dt_sp_create: out of memory");
        if (hipMemcpyAsync(sp->val64, val64, (size_t)nnz * 8, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            return fail(DT_EHIP, "dt_sp_create: copy failed");
        if (dt_convert(ctx, sp->val, DT_F32, sp->val64, DT_F64, (size_t)nnz) != DT_OK)
            return fail(DT_EHIP, "dt_sp_create: conversion failed");
    } else if (hipMemcpyAsync(sp->val, val, (size_t)nnz * 4, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
        return fail(DT_EHIP, "dt_sp_create: copy failed");
    }
    for (int n = 0; n < ndim; n++) {
        dt_sp_mode &m = sp->mode[n];
        const int64_t I = sp->dims[n];
        m.perm = (int32_t *)sp_alloc(sp, (size_t)nnz * 4);
        m.row = (int32_t *)sp_alloc(sp, (size_t)nnz * 4);
        m.rowptr = (int32_t *)sp_alloc(sp, (size_t)(I + 1) * 4);
        m.slot = (int32_t *)sp_alloc(sp, (size_t)(I + 1) * 4);
        if (!m.perm || !m.row || !m.rowptr || !m.slot) return fail(DT_ENOMEM, "dt_sp_create: out of memory");
        if (ndim == 3) {
            m.rec = (intx4 *)sp_alloc(sp, (size_t)nnz * 16);
            if (!m.rec) return fail(DT_ENOMEM, "dt_sp_create: out of memory");
        }
        st = sort_mode_internal(ctx, nnz, sp->idx[n], I, m.perm, m.row, m.rowptr);
        if (st != DT_OK) return fail(st, "dt_sp_create: mode sort failed");
        // compact numbering of the non-empty rows
        int32_t *f = (int32_t *)dt_ws(ctx, WS_SPARSE1, (size_t)(I + 1) * 4 + 256);
        if (!f) return fail(DT_ENOMEM, "dt_sp_create: out of memory");
        hipLaunchKernelGGL(nonempty_flags, sp_grid(I), dim3(256), 0, ctx->stream, m.rowptr, I, f);
        int32_t *stmp = (int32_t *)sort_tmp(ctx, dt_scan_tmp_bytes(I));
        if (!stmp || dt_exclusive_scan_i32(ctx, f, m.slot, I, stmp) != DT_OK)
            return fail(DT_EHIP, "dt_sp_create: scan failed");
        int32_t lastv[2];
        if (hipMemcpyAsync(&lastv[0], m.slot + I - 1, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(&lastv[1], f + I - 1, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return fail(DT_EHIP, "dt_sp_create: readback failed");
        m.nrows = (int64_t)lastv[0] + lastv[1];
        m.rows = (int32_t *)sp_alloc(sp, (size_t)std::max<int64_t>(m.nrows, 1) * 4);
        if (!m.rows) return fail(DT_ENOMEM, "dt_sp_create: out of memory");
        hipLaunchKernelGGL(rows_list, sp_grid(I), dim3(256), 0, ctx->stream, f, m.slot, I, m.rows);
        if (ndim == 3) {
            // TTMc records: the other modes' indices and the value in this mode's order
            const int a = n == 0 ? 1 : 0, b = n == 2 ? 1 : 2;
            int32_t *tia = (int32_t *)dt_ws(ctx, WS_SPARSE0, (size_t)nnz * 12 + 256);
            if (!tia) return fail(DT_ENOMEM, "dt_sp_create: out of memory");
            int32_t *tib = tia + nnz;
            float *tval = (float *)(tib + nnz);
            hipLaunchKernelGGL(gather_by<int32_t>, sp_grid(nnz), dim3(256), 0, ctx->stream, sp->idx[a], m.perm, nnz, tia);
            hipLaunchKernelGGL(gather_by<int32_t>, sp_grid(nnz), dim3(256), 0, ctx->stream, sp->idx[b], m.perm, nnz, tib);
            hipLaunchKernelGGL(gather_by<float>, sp_grid(nnz), dim3(256), 0, ctx->stream, sp->val, m.perm, nnz, tval);
            hipLaunchKernelGGL(build_rec, sp_grid(nnz), dim3(256), 0, ctx->stream, m.row, tia, tib, tval, m.slot, nnz,
                               m.rec);
            // inside each row, order the records by the index in mode a (the
            // larger of the two gathered factors): a long row's U_a gathers
            // then sweep U_a in order instead of jumping across it.  Rows keep
            // their place in the stream, so the mode permutation and the row
            // pointers are unchanged.
            st = sort_rows_by_a(ctx, nnz, bits_for(sp->dims[a]), bits_for(std::max<int64_t>(m.nrows, 1)), m.rec);
            if (st != DT_OK) return fail(st, "dt_sp_create: record sort failed");
        }
        if (hipGetLastError() != hipSuccess) return fail(DT_EHIP, "dt_sp_create: launch failed");
    }
    // ||A||^2 in FP64 (of the FP64 values when the tensor has them)
    if (sp->val64) {
        if (dt_sumsq_dev(ctx, sp->val64, DT_F64, (size_t)nnz, &sp->normA2) != DT_OK)
            return fail(DT_EHIP, "dt_sp_create: norm failed");
        *out = sp;
        return DT_OK;
    }
    double *part = (double *)dt_ws(ctx, WS_SUMSQ, 1024 * 8);
    if (!part) return fail(DT_ENOMEM, "dt_sp_create: out of memory");
    hipLaunchKernelGGL(sumsq_f32_sp, dim3(1024), dim3(256), 0, ctx->stream, sp->val, nnz, part);
    std::vector<double> hp(1024);
    if (hipMemcpyAsync(hp.data(), part, 1024 * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail(DT_EHIP, "dt_sp_create: readback failed");
    for (double v : hp) sp->normA2 += v;
    *out = sp;
    return DT_OK;
}

extern "C" int dt_sp_create(dt_ctx *ctx, int ndim, const size_t *dims, int64_t nnz, const int32_t *const *idx,
                            const float *val, dt_sp **out) {
    DT_CHECK_ARG(val, "dt_sp_create: null values");
    return sp_create_impl(ctx, ndim, dims, nnz, idx, val, nullptr, out);
}

extern "C" int dt_sp_create_f64(dt_ctx *ctx, int ndim, const size_t *dims, int64_t nnz, const int32_t *const *idx,
                                const double *val, dt_sp **out) {
    DT_CHECK_ARG(val, "dt_sp_create_f64: null values");
    return sp_create_impl(ctx, ndim, dims, nnz, idx, nullptr, val, out);
}

extern "C" int dt_sp_info(const dt_sp *sp, int *slice_mode, int64_t *nnz, double *norm) {
    if (!sp) return DT_EINVAL;
    if (slice_mode) *slice_mode = sp->slice_mode;
    if (nnz) *nnz = sp->nnz;
    if (norm) *norm = std::sqrt(sp->normA2);
    return DT_OK;
}

extern "C" int dt_sp_mode_arrays(const dt_sp *sp, int n, const int32_t **perm, const int32_t **rowptr,
                                 int64_t *nonempty_rows) {
    if (!sp || n < 0 || n >= sp->ndim) return DT_EINVAL;
    if (perm) *perm = sp->mode[n].perm;
    if (rowptr) *rowptr = sp->mode[n].rowptr;
    if (nonempty_rows) *nonempty_rows = sp->mode[n].nrows;
    return DT_OK;
}

// ---- TTMc driver ---------------------------------------------------------------

// L = Y_cᵀ U_n (256 x 16 FP64) from the compact rows of mode n; Un row-major
// FP32 I_n x 16.
static int sp_ytu_dev(dt_ctx *ctx, dt_sp *sp, int n, const float *Un, const float *Yc, double *L) {
    const dt_sp_mode &m = sp->mode[n];
    // at most 8192 waves of at least 16 rows; one partial per 4-wave workgroup
    const int64_t nw = std::min<int64_t>(8192, std::max<int64_t>(1, (m.nrows + 15) / 16));
    const int64_t rpw = (m.nrows + nw - 1) / nw;
    const int64_t ng = std::max<int64_t>(1, ((m.nrows + rpw - 1) / rpw + 3) / 4);
    float *Lwg = (float *)dt_ws(ctx, WS_SPARSE1, (size_t)ng * SP_E * 16 * 4 + 256);
    if (!Lwg) return DT_ENOMEM;
    DT_LAUNCH("sp_ytu", sp_ytu, dim3((unsigned)ng), dim3(256), 0, Yc, m.rows, m.nrows, Un, rpw, Lwg);
    DT_LAUNCH("sp_lreduce", sp_lreduce, dim3(SP_E * 16 / 16), dim3(256), 0, (const float *)Lwg, (int)ng, L);
    return DT_OK;
}

// Yc (nrows x 256) and, with L, L = Y_cᵀ U_n (256 x 16 FP64) for mode n of a
// 3-way tensor with ranks 16 in the other modes.  Ur[m]: row-major FP32 I_m x 16.
static int sp_ttmc_dev(dt_ctx *ctx, dt_sp *sp, int n, const float *const *Ur, float *Yc, double *L) {
    DT_CHECK_ARG(sp->ndim == 3, "sparse TTMc: 3-way tensors only (got %d)", sp->ndim);
    const dt_sp_mode &m = sp->mode[n];
    const int a = n == 0 ? 1 : 0, b = n == 2 ? 1 : 2;
    SpModeArgs args{m.row, m.rec, m.rowptr, m.slot, sp->nnz};
    // 32-bit gather offsets when both gathered factors (rows x 64 B) are below 4 GiB
    const bool off32 = sp->dims[a] < ((int64_t)1 << 26) && sp->dims[b] < ((int64_t)1 << 26);
    const int64_t wgs = dt_resident_wgs(ctx, sp_ttmc16<true>, 256);
    if (wgs <= 0) return dt_fail(ctx, DT_EHIP, "sparse TTMc: occupancy query failed");
    // equal contiguous ranges, multiples of 64, about eight per resident wave
    // (the hardware hands a range to whichever wave finishes first, which
    // evens out ranges of different cost: short-row stretches take longer)
    const int64_t maxw = wgs * 4 * 8;
    int64_t per_wave = ((sp->nnz + maxw - 1) / maxw + 63) / 64 * 64;
    if (per_wave < 64) per_wave = 64;
    const int64_t nwaves = (sp->nnz + per_wave - 1) / per_wave;
    // carry slots (2 per range), the chunk sums of long chains (2 per range),
    // the list of ranges with fixup work and its length
    float *carry = (float *)dt_ws(ctx, WS_SPARSE0, (size_t)nwaves * (4 * SP_E * 4 + 4) + 512);
    if (!carry) return DT_ENOMEM;
    float *csum = carry + (size_t)nwaves * 2 * SP_E;
    int *fix_list = (int *)(csum + (size_t)nwaves * 2 * SP_E);
    int *fix_count = fix_list + nwaves;
    DT_TRY(dt_zero(ctx, fix_count, sizeof(int)));
    static const char *names[3] = {"sp_ttmc_m0", "sp_ttmc_m1", "sp_ttmc_m2"};
    if (off32) {
        DT_LAUNCH(names[n], sp_ttmc16<true>, dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, args, Ur[a], Ur[b], Yc,
                  carry, per_wave, nwaves, fix_list, fix_count);
    } else {
        DT_LAUNCH(names[n], sp_ttmc16<false>, dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, args, Ur[a], Ur[b],
                  Yc, carry, per_wave, nwaves, fix_list, fix_count);
    }
    // one-wave workgroups, as many as ranges (the list has at most one entry
    // per range, so each wave takes at most one)
    const unsigned nfix = (unsigned)nwaves;
    DT_LAUNCH("sp_fixup", sp_fixup, dim3(nfix), dim3(64), 0, args, carry, per_wave, csum, Yc, (const int *)fix_list,
              (const int *)fix_count);
    DT_LAUNCH("sp_fixup", sp_fixup_chunks, dim3(nfix), dim3(64), 0, args, (const float *)csum, per_wave, Yc,
              (const int *)fix_list, (const int *)fix_count);
    if (L) DT_TRY(sp_ytu_dev(ctx, sp, n, Ur[n], Yc, L));
    return DT_OK;
}

extern "C" int dt_sp_ttmc(dt_ctx *ctx, dt_sp *sp, int n, const float *const *Ur, const size_t *ranks, float *Y) {
    DT_CHECK_ARG(ctx && sp && Ur && ranks && Y, "dt_sp_ttmc: null argument");
    DT_CHECK_ARG(n >= 0 && n < sp->ndim, "dt_sp_ttmc: bad mode %d", n);
    DT_CHECK_ARG(sp->ndim == 3, "dt_sp_ttmc: 3-way tensors only");
    for (int m = 0; m < 3; m++)
        if (m != n) DT_CHECK_ARG(ranks[m] == 16 && Ur[m], "dt_sp_ttmc: ranks of the other modes must be 16");
    const dt_sp_mode &md = sp->mode[n];
    float *Yc = (float *)dt_ws(ctx, WS_SPARSE2, (size_t)std::max<int64_t>(md.nrows, 1) * SP_E * 4);
    if (!Yc) return DT_ENOMEM;
    DT_TRY(sp_ttmc_dev(ctx, sp, n, Ur, Yc, nullptr));
    DT_TRY(dt_zero(ctx, Y, (size_t)sp->dims[n] * SP_E * 4));
    DT_LAUNCH("sp_expand", sp_expand, sp_grid(md.nrows * SP_E), dim3(256), 0, Yc, md.rows, md.nrows, Y);
    return DT_OK;
}

// ---- sparse HOOI -----------------------------------------------------------------
// A session holds the factors and the buffers of one decomposition; every
// sweep runs Sp_TTMC_ALS_ (sp_hooi_als.tpp:48-65) for each mode in order and
// forms the core from the last mode.  Two paths:
//   fast    : 3-way, every rank <= 16, FP32 values -- the MFMA TTMc pipeline
//             above, factors as FP32 rows padded to 16;
//   generic : any other shape (dt_sparse_nway.hip), in FP64 for a tensor with
//             FP64 values, else FP32; factor update W = Y_(n) L, then
//             shifted Cholesky-QR3 (dt_orth_dev).
// No step waits for the host inside a sweep; the sweep reads back the core's
// norm, each mode's ||L|| and the breakdown flag once.
struct dt_sp_hooi {
    dt_ctx *ctx = nullptr;
    dt_sp *sp = nullptr;
    int N = 0;
    int R[DT_MAX_NDIM] = {};
    bool fast = false;
    int T = DT_F32;               // generic path: element type
    std::vector<void *> mine;
    double *U64[DT_MAX_NDIM] = {};
    void *Urow[DT_MAX_NDIM] = {};  // fast: FP32 I x 16 rows; generic: T rows I x R
    void *Yc = nullptr;
    double *L = nullptr, *G64 = nullptr;
    double *part = nullptr, *Gs = nullptr, *work = nullptr;
    int *dflag = nullptr;
    double *scal = nullptr;       // [0] = ||G||, [1 + n] = ||L|| of mode n
    int64_t Klast = 0;            // core: Klast x R[N-1]
    int sweeps = 0;
    double total_ms = 0.0, res = 0.0;
};

constexpr int kSpMaxWg = 2048;  // fast factor update: workgroups (Gram partials)

static void *sh_alloc(dt_sp_hooi *h, size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    h->mine.push_back(p);
    return p;
}

// fast path: U_n <- orth(Y_(n) L): sp_tl_gram, then three Cholesky-QR passes
// (the first shifted) whose Grams come from the previous pass's partials; the
// last pass also writes the FP32 view
static int sh_fast_update(dt_sp_hooi *h, int n) {
    dt_ctx *ctx = h->ctx;
    dt_sp *sp = h->sp;
    const int64_t I = sp->dims[n];
    const int Rn = h->R[n];
    const int64_t ntiles = (I + 15) / 16;
    const int64_t tpw = std::max<int64_t>(1, (ntiles + 4 * kSpMaxWg - 1) / (4 * kSpMaxWg));
    const int ng = (int)(((ntiles + tpw - 1) / tpw + 3) / 4);
    const dt_sp_mode &md = sp->mode[n];
    DT_LAUNCH("sp_tl_gram", sp_tl_gram, dim3((unsigned)ng), dim3(256), 0, (const float *)h->Yc, md.rowptr, md.slot, I,
              Rn, (const double *)h->L, tpw, h->U64[n], h->part);
    auto rmul_gram = sp_rmul<true, false>;
    auto rmul_view = sp_rmul<false, true>;
    for (int pass = 0; pass < 3; pass++) {
        DT_TRY(dt_part_reduce(ctx, h->part, ng, Rn * Rn, h->Gs));
        const int shifted = pass == 0 ? 1 : 0;
        if (pass < 2) {
            DT_LAUNCH("sp_rmul", rmul_gram, dim3((unsigned)ng), dim3(256), 0, h->U64[n], I, Rn, (const double *)h->Gs,
                      shifted, h->dflag, tpw, h->part, (float *)nullptr);
        } else {
            DT_LAUNCH("sp_rmul", rmul_view, dim3((unsigned)ng), dim3(256), 0, h->U64[n], I, Rn, (const double *)h->Gs,
                      shifted, h->dflag, tpw, (double *)nullptr, (float *)h->Urow[n]);
        }
    }
    return DT_OK;
}

static int sh_views(dt_sp_hooi *h, int m) {
    dt_ctx *ctx = h->ctx;
    if (h->fast) {
        DT_LAUNCH("sp_views", rows16_view, sp_grid(h->sp->dims[m]), dim3(256), 0, h->U64[m], h->sp->dims[m], h->R[m],
                  (float *)h->Urow[m]);
        return DT_OK;
    }
    return dt_spg_rows_view(ctx, h->U64[m], h->sp->dims[m], h->R[m], h->Urow[m], h->T);
}

extern "C" int dt_sp_hooi_destroy(dt_sp_hooi *h) {
    if (!h) return DT_OK;
    (void)hipStreamSynchronize(h->ctx->stream);
    for (void *p : h->mine) (void)hipFree(p);
    delete h;
    return DT_OK;
}

extern "C" int dt_sp_hooi_create(dt_ctx *ctx, dt_sp *sp, const size_t *ranks, uint64_t seed,
                                 const double *const *U_init, dt_sp_hooi **out) {
    DT_CHECK_ARG(ctx && sp && ranks && out, "dt_sp_hooi_create: null argument");
    *out = nullptr;
    const int N = sp->ndim;
    bool small = N == 3 && sp->val64 == nullptr;
    int64_t K[DT_MAX_NDIM];
    for (int m = 0; m < N; m++) {
        DT_CHECK_ARG(ranks[m] >= 1 && (int64_t)ranks[m] <= sp->dims[m] && ranks[m] <= (size_t)DT_SPG_RMAX,
                     "dt_sp_hooi_create: rank %zu of mode %d not in [1, min(%d, I = %lld)]", ranks[m], m, DT_SPG_RMAX,
                     (long long)sp->dims[m]);
        small = small && ranks[m] <= 16;
    }
    int64_t maxK = 1;
    for (int n = 0; n < N; n++) {
        K[n] = 1;
        for (int m = 0; m < N; m++)
            if (m != n) K[n] *= (int64_t)ranks[m];
        maxK = std::max(maxK, K[n]);
    }
    for (int n = 0; n < N; n++)
        DT_CHECK_ARG((int64_t)ranks[n] <= K[n],
                     "dt_sp_hooi_create: rank %zu of mode %d above the product %lld of the other ranks (Y_(n) has "
                     "that many columns, so its leading subspace has at most that dimension)",
                     ranks[n], n, (long long)K[n]);
    DT_CHECK_ARG(small || maxK <= DT_SPG_KMAX,
                 "dt_sp_hooi_create: product of the other modes' ranks %lld above %d", (long long)maxK, DT_SPG_KMAX);
    dt_sp_hooi *h = new dt_sp_hooi();
    h->ctx = ctx;
    h->sp = sp;
    h->N = N;
    h->fast = small;
    h->T = sp->val64 ? DT_F64 : DT_F32;
    for (int m = 0; m < N; m++) h->R[m] = (int)ranks[m];
    h->Klast = h->fast ? SP_E : K[N - 1];
    const size_t tsz = h->T == DT_F64 ? 8 : 4;
    int64_t maxrows = 0, maxI = 0, maxR = 0;
    for (int m = 0; m < N; m++) {
        maxrows = std::max(maxrows, sp->mode[m].nrows);
        maxI = std::max(maxI, sp->dims[m]);
        maxR = std::max<int64_t>(maxR, h->R[m]);
    }
    bool ok = true;
    for (int m = 0; m < N && ok; m++) {
        h->U64[m] = (double *)sh_alloc(h, (size_t)sp->dims[m] * h->R[m] * 8);
        h->Urow[m] = sh_alloc(h, h->fast ? (size_t)sp->dims[m] * 16 * 4 : (size_t)sp->dims[m] * h->R[m] * tsz);
        ok = h->U64[m] && h->Urow[m];
    }
    const int64_t ycw = h->fast ? SP_E : maxK;
    if (ok) ok = (h->Yc = sh_alloc(h, (size_t)std::max<int64_t>(maxrows, 1) * ycw * (h->fast ? 4 : tsz))) != nullptr;
    const size_t lsz = h->fast ? (size_t)SP_E * 16 : (size_t)maxK * maxR;
    if (ok) ok = (h->L = (double *)sh_alloc(h, lsz * 8)) != nullptr;
    if (ok) ok = (h->G64 = (double *)sh_alloc(h, lsz * 8)) != nullptr;
    if (ok && h->fast) ok = (h->part = (double *)sh_alloc(h, (size_t)kSpMaxWg * 256 * 8)) != nullptr;
    if (ok) ok = (h->Gs = (double *)sh_alloc(h, 256 * 8)) != nullptr;
    if (ok && !h->fast) ok = (h->work = (double *)sh_alloc(h, (size_t)(2 * maxR * maxR + maxI * maxR + 64) * 8)) != nullptr;
    if (ok) ok = (h->dflag = (int *)sh_alloc(h, 256)) != nullptr;
    if (ok) ok = (h->scal = (double *)sh_alloc(h, (DT_MAX_NDIM + 1) * 8)) != nullptr;
    if (!ok) {
        dt_sp_hooi_destroy(h);
        return dt_fail(ctx, DT_ENOMEM, "dt_sp_hooi_create: out of memory");
    }
    int st = dt_zero(ctx, h->dflag, sizeof(int));
    if (st == DT_OK) st = dt_zero(ctx, h->scal, (DT_MAX_NDIM + 1) * 8);
    if (h->fast && st == DT_OK)
        for (int m = 0; m < N && st == DT_OK; m++) st = dt_zero(ctx, h->Urow[m], (size_t)sp->dims[m] * 16 * 4);
    // start: the caller's factors, or random orthonormal (Philox stream 200 + m);
    // QR with positive R diagonal, the basis orth2's Gram-Schmidt gives
    for (int m = 0; m < N && st == DT_OK; m++) {
        const size_t sz = (size_t)sp->dims[m] * h->R[m];
        if (U_init && U_init[m]) {
            if (hipMemcpyAsync(h->U64[m], U_init[m], sz * 8, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
                st = dt_fail(ctx, DT_EHIP, "dt_sp_hooi_create: copy failed");
        } else {
            st = dt_fill_normal_f64(ctx, h->U64[m], sz, seed, 200 + (uint64_t)m);
        }
        if (st == DT_OK) st = dt_qr_f64(ctx, h->U64[m], nullptr, h->U64[m], (size_t)sp->dims[m], (size_t)h->R[m]);
        if (st == DT_OK) st = sh_views(h, m);
    }
    if (st != DT_OK) {
        dt_sp_hooi_destroy(h);
        return st;
    }
    *out = h;
    return DT_OK;
}

extern "C" int dt_sp_hooi_sweep(dt_sp_hooi *h, double *L_norms, double *residual) {
    if (!h) return DT_EINVAL;
    dt_ctx *ctx = h->ctx;
    dt_sp *sp = h->sp;
    const int N = h->N;
    hipEvent_t e0 = dt_get_event(ctx), e1 = dt_get_event(ctx);
    (void)hipEventRecord(e0, ctx->stream);
    int st = DT_OK;
    for (int n = 0; n < N && st == DT_OK; n++) {
        if (h->fast) {
            const float *Ur[3] = {(const float *)h->Urow[0], (const float *)h->Urow[1], (const float *)h->Urow[2]};
            st = sp_ttmc_dev(ctx, sp, n, Ur, (float *)h->Yc, h->L);  // Y_c and L = Y_cᵀ U_n
            if (st == DT_OK) st = dt_spg_norm(ctx, h->L, (int64_t)SP_E * 16, h->scal + 1 + n);
            if (st == DT_OK) st = sh_fast_update(h, n);              // U_n <- orth(Y_(n) L)
        } else {
            int64_t K = 1;
            for (int m = 0; m < N; m++)
                if (m != n) K *= h->R[m];
            st = dt_spg_ttmc(ctx, sp, n, h->Urow, h->T, h->R, h->Yc);
            if (st == DT_OK) st = dt_spg_ytu(ctx, sp, n, h->Urow[n], h->T, h->R, h->Yc, h->L);
            if (st == DT_OK) st = dt_spg_norm(ctx, h->L, K * h->R[n], h->scal + 1 + n);
            if (st == DT_OK) st = dt_spg_yl(ctx, sp, n, h->T, h->R, h->Yc, h->L, h->U64[n]);
            if (st == DT_OK) st = dt_orth_dev(ctx, h->U64[n], (size_t)sp->dims[n], (size_t)h->R[n], h->work, h->dflag);
            if (st == DT_OK) st = sh_views(h, n);
        }
    }
    // core = Y_(N-1)ᵀ U_{N-1} with the updated U_{N-1}: Y_(N-1) does not depend
    // on U_{N-1}, so the last mode's Y_c is reused
    if (st == DT_OK) {
        if (h->fast)
            st = sp_ytu_dev(ctx, sp, N - 1, (const float *)h->Urow[N - 1], (const float *)h->Yc, h->G64);
        else
            st = dt_spg_ytu(ctx, sp, N - 1, h->Urow[N - 1], h->T, h->R, h->Yc, h->G64);
    }
    if (st == DT_OK) st = dt_spg_norm(ctx, h->G64, h->Klast * (h->fast ? 16 : h->R[N - 1]), h->scal);
    (void)hipEventRecord(e1, ctx->stream);
    double sc[DT_MAX_NDIM + 1];
    if (st == DT_OK &&
        (hipMemcpyAsync(sc, h->scal, sizeof(sc), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
         hipMemcpyAsync(ctx->pinned, h->dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
         hipStreamSynchronize(ctx->stream) != hipSuccess))
        st = dt_fail(ctx, DT_EHIP, "dt_sp_hooi_sweep: readback failed");
    float ms = 0.f;
    if (st == DT_OK) (void)hipEventElapsedTime(&ms, e0, e1);
    ctx->event_pool.push_back(e0);
    ctx->event_pool.push_back(e1);
    if (st != DT_OK) return st;
    if (*(const int *)ctx->pinned != 0)
        return dt_fail(ctx, DT_ENUMERIC,
                       "dt_sp_hooi_sweep: Cholesky breakdown in the factor update of sweep %d (Y_(n) L rank-deficient)",
                       h->sweeps);
    h->total_ms += ms;
    h->sweeps++;
    const double g2 = sc[0] * sc[0];
    h->res = std::sqrt(std::max(0.0, 1.0 - g2 / sp->normA2));
    if (L_norms)
        for (int n = 0; n < N; n++) L_norms[n] = sc[1 + n];
    if (residual) *residual = h->res;
    return DT_OK;
}

extern "C" int dt_sp_hooi_get(dt_sp_hooi *h, double *const *U_out, double *G_out, double *sweeps_ms) {
    if (!h) return DT_EINVAL;
    dt_ctx *ctx = h->ctx;
    dt_sp *sp = h->sp;
    const int N = h->N;
    for (int m = 0; m < N; m++)
        if (U_out && U_out[m])
            DT_HIP(hipMemcpyAsync(U_out[m], h->U64[m], (size_t)sp->dims[m] * h->R[m] * 8, hipMemcpyDeviceToDevice,
                                  ctx->stream));
    if (G_out) {
        DT_CHECK_ARG(h->sweeps > 0, "dt_sp_hooi_get: the core needs a sweep");
        if (h->fast) {
            // G (R0 x R1 x R2, column-major): G[r0 + R0 r1 + R0 R1 r2] = L[(r0 + 16 r1) * 16 + r2]
            std::vector<double> g(SP_E * 16), gc((size_t)h->R[0] * h->R[1] * h->R[2]);
            DT_HIP(hipMemcpyAsync(g.data(), h->G64, g.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
            DT_HIP(hipStreamSynchronize(ctx->stream));
            for (int r2 = 0; r2 < h->R[2]; r2++)
                for (int r1 = 0; r1 < h->R[1]; r1++)
                    for (int r0 = 0; r0 < h->R[0]; r0++)
                        gc[(size_t)(r0 + h->R[0] * (r1 + h->R[1] * r2))] = g[(size_t)((r0 + 16 * r1) * 16 + r2)];
            DT_HIP(hipMemcpyAsync(G_out, gc.data(), gc.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        } else {
            // L of the last mode (K x R_{N-1}, column-major, K in the other modes'
            // order with the first fastest) is the column-major core
            DT_HIP(hipMemcpyAsync(G_out, h->G64, (size_t)h->Klast * h->R[N - 1] * 8, hipMemcpyDeviceToDevice,
                                  ctx->stream));
        }
    }
    if (sweeps_ms) *sweeps_ms = h->total_ms;
    DT_HIP(hipStreamSynchronize(ctx->stream));
    return DT_OK;
}

extern "C" int dt_hooi_sparse(dt_ctx *ctx, dt_sp *sp, const size_t *ranks, int max_iter, uint64_t seed,
                              const double *const *U_init, double *const *U_out, double *G_out,
                              double *residual_per_sweep, double *residual, double *sweeps_ms) {
    DT_CHECK_ARG(ctx && sp && ranks && residual, "dt_hooi_sparse: null argument");
    DT_CHECK_ARG(max_iter >= 1, "dt_hooi_sparse: max_iter >= 1");
    dt_sp_hooi *h = nullptr;
    DT_TRY(dt_sp_hooi_create(ctx, sp, ranks, seed, U_init, &h));
    int st = DT_OK;
    for (int it = 0; it < max_iter && st == DT_OK; it++) {
        double r = 0.0;
        st = dt_sp_hooi_sweep(h, nullptr, &r);
        if (st == DT_OK && residual_per_sweep) residual_per_sweep[it] = r;
        if (st == DT_OK) *residual = r;
    }
    if (st == DT_OK) st = dt_sp_hooi_get(h, U_out, G_out, sweeps_ms);
    dt_sp_hooi_destroy(h);
    return st;
}
