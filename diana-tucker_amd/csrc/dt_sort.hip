// Stable LSD radix sort of (key, int32 payload) pairs and an exclusive prefix
// sum of int32 counts, hand-written for the sparse tensor's index work.
//
// The reference's SpTensor constructor sorts the nonzeros stably by the slice
// mode with a counting sort: histogram, exclusive scan, scatter in input order
// (include/sp_tensor.hpp:117-147).  One pass of this sort is that counting
// sort over 8 bits of the key; ceil(nbits / 8) passes from the lowest digit
// give the stable order of the whole key, so a sort by the mode index alone
// produces exactly the constructor's permutation.
//
// Per pass, over tiles of RS_TILE = 4096 items (256 threads x 16 rounds):
//   rs_hist    : per tile, the count of each of the 256 digits (LDS atomics),
//                stored digit-major: hist[digit * ntiles + tile];
//   scan       : exclusive prefix sum of hist, so digit d of tile t starts
//                after every item of a smaller digit and after digit d of every
//                earlier tile (stable across tiles);
//   rs_scatter : each tile walks its items in input order, 256 per round; in
//                a round every wave finds the lanes that share its digit with
//                eight ballots (one per digit bit), a lane's rank among them is
//                the count of such lanes below it, the waves' group sizes go
//                through LDS, and the tile's running offset per digit grows by
//                the round's totals -- so equal digits keep their input order.
#include "dt_internal.hpp"

#include <algorithm>

namespace {

constexpr int RS_THREADS = 256, RS_ROUNDS = 16, RS_TILE = RS_THREADS * RS_ROUNDS, RS_BINS = 256;
constexpr int SC_TILE = 4096;  // scan: items per workgroup (16 per thread)

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void rs_hist(const K *__restrict__ keys, int64_t n, int shift,
                                                      int32_t *__restrict__ hist, int64_t ntiles) {
    __shared__ int32_t cnt[RS_BINS];
    const int tid = threadIdx.x;
    cnt[tid] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int q = 0; q < RS_ROUNDS; q++) {
        const int64_t i = b0 + (int64_t)q * RS_THREADS + tid;
        if (i < n) atomicAdd(&cnt[(unsigned)(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[(int64_t)tid * ntiles + blockIdx.x] = cnt[tid];
}

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void rs_scatter(const K *__restrict__ kin, const int32_t *__restrict__ vin,
                                                         K *__restrict__ kout, int32_t *__restrict__ vout, int64_t n,
                                                         int shift, const int32_t *__restrict__ off, int64_t ntiles) {
    __shared__ int32_t base[RS_BINS];
    __shared__ int32_t wcnt[RS_THREADS / 64][RS_BINS];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    base[tid] = off[(int64_t)tid * ntiles + blockIdx.x];
#pragma unroll
    for (int u = 0; u < RS_THREADS / 64; u++) wcnt[u][tid] = 0;
    __syncthreads();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t b0 = (int64_t)blockIdx.x * RS_TILE;
    for (int q = 0; q < RS_ROUNDS; q++) {
        const int64_t i = b0 + (int64_t)q * RS_THREADS + tid;
        if (b0 + (int64_t)q * RS_THREADS >= n) break;  // uniform: no item of this round is valid
        const bool ok = i < n;
        const K k = ok ? kin[i] : (K)0;
        const int32_t v = ok ? vin[i] : 0;
        const unsigned d = (unsigned)(k >> shift) & 255u;
        // lanes of this wave holding the same digit (valid lanes only)
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        if (ok && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (ok) {
            int32_t pos = base[d] + rank;
            for (int u = 0; u < w; u++) pos += wcnt[u][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        int32_t add = 0;
#pragma unroll
        for (int u = 0; u < RS_THREADS / 64; u++) {
            add += wcnt[u][tid];
            wcnt[u][tid] = 0;
        }
        base[tid] += add;
        __syncthreads();
    }
}

// ---- exclusive prefix sum of int32 ------------------------------------------
__device__ __forceinline__ int32_t wave_incl_scan(int32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// exclusive scan of one value per thread over the workgroup (blockDim.x a
// multiple of 64, at most 1024); *total receives the sum
__device__ int32_t block_excl_scan(int32_t x, int32_t *total) {
    __shared__ int32_t ws[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int32_t inc = wave_incl_scan(x);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    int32_t before = 0, tot = 0;
    for (int u = 0; u < nw; u++) {
        if (u < w) before += ws[u];
        tot += ws[u];
    }
    __syncthreads();
    *total = tot;
    return before + inc - x;
}

__global__ __launch_bounds__(256) void scan_tile_sums(const int32_t *__restrict__ in, int64_t n,
                                                      int32_t *__restrict__ sums) {
    const int64_t b0 = (int64_t)blockIdx.x * SC_TILE;
    int32_t s = 0;
    for (int q = 0; q < SC_TILE / 256; q++) {
        const int64_t i = b0 + (int64_t)q * 256 + threadIdx.x;
        if (i < n) s += in[i];
    }
    int32_t tot;
    (void)block_excl_scan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// one workgroup: exclusive scan of the nt tile sums in place
__global__ __launch_bounds__(1024) void scan_sums(int32_t *__restrict__ sums, int64_t nt) {
    const int64_t per = (nt + 1023) / 1024;
    const int64_t s0 = (int64_t)threadIdx.x * per, s1 = std::min<int64_t>(s0 + per, nt);
    int32_t local = 0;
    for (int64_t i = s0; i < s1; i++) local += sums[i];
    int32_t tot;
    int32_t run = block_excl_scan(local, &tot);
    for (int64_t i = s0; i < s1; i++) {
        const int32_t x = sums[i];
        sums[i] = run;
        run += x;
    }
}

// each thread scans 16 consecutive items of the tile, offset by the threads
// before it and the tile's exclusive prefix; in and out may alias
__global__ __launch_bounds__(256) void scan_tiles(const int32_t *in, int64_t n, const int32_t *__restrict__ sums,
                                                  int32_t *out) {
    constexpr int PER = SC_TILE / 256;
    const int64_t i0 = (int64_t)blockIdx.x * SC_TILE + (int64_t)threadIdx.x * PER;
    int32_t x[PER];
    int32_t s = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        x[q] = i0 + q < n ? in[i0 + q] : 0;
        s += x[q];
    }
    int32_t tot;
    int32_t run = block_excl_scan(s, &tot) + sums[blockIdx.x];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        if (i0 + q < n) out[i0 + q] = run;
        run += x[q];
    }
}

template <typename K>
int radix_sort_pairs(dt_ctx *ctx, const K *kin, K *kout, const int32_t *vin, int32_t *vout, int64_t n, int nbits,
                     void *tmp) {
    if (n <= 0) return DT_OK;
    const int64_t ntiles = ceil_div(n, RS_TILE);
    const int64_t nh = ntiles * RS_BINS;
    // tmp: alternate keys | alternate payloads | histogram | scan sums
    char *t = (char *)tmp;
    K *kalt = (K *)t;
    t += ((size_t)n * sizeof(K) + 255) & ~(size_t)255;
    int32_t *valt = (int32_t *)t;
    t += ((size_t)n * 4 + 255) & ~(size_t)255;
    int32_t *hist = (int32_t *)t;
    t += ((size_t)nh * 4 + 255) & ~(size_t)255;
    int32_t *sums = (int32_t *)t;
    const int passes = std::max(1, (nbits + 7) / 8);
    // ping-pong so that the last pass writes kout / vout
    const K *ks = kin;
    const int32_t *vs = vin;
    for (int p = 0; p < passes; p++) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        K *kd = to_out ? kout : kalt;
        int32_t *vd = to_out ? vout : valt;
        const int shift = 8 * p;
        DT_LAUNCH("radix_sort", rs_hist<K>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, ks, n, shift, hist, ntiles);
        DT_TRY(dt_exclusive_scan_i32(ctx, hist, hist, nh, sums));
        DT_LAUNCH("radix_sort", rs_scatter<K>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, ks, vs, kd, vd, n, shift,
                  (const int32_t *)hist, ntiles);
        ks = kd;
        vs = vd;
    }
    return DT_OK;
}

}  // namespace

size_t dt_scan_tmp_bytes(int64_t n) { return ((size_t)ceil_div(std::max<int64_t>(n, 1), SC_TILE) * 4 + 255) & ~(size_t)255; }

int dt_exclusive_scan_i32(dt_ctx *ctx, const int32_t *in, int32_t *out, int64_t n, int32_t *tmp) {
    if (n <= 0) return DT_OK;
    const int64_t nt = ceil_div(n, SC_TILE);
    DT_LAUNCH("scan", scan_tile_sums, dim3((unsigned)nt), dim3(256), 0, in, n, tmp);
    DT_LAUNCH("scan", scan_sums, dim3(1), dim3(1024), 0, tmp, nt);
    DT_LAUNCH("scan", scan_tiles, dim3((unsigned)nt), dim3(256), 0, in, n, (const int32_t *)tmp, out);
    return DT_OK;
}

size_t dt_radix_sort_tmp_bytes(int64_t n, int key_bytes) {
    const int64_t ntiles = ceil_div(std::max<int64_t>(n, 1), RS_TILE);
    const int64_t nh = ntiles * RS_BINS;
    return (((size_t)n * key_bytes + 255) & ~(size_t)255) + (((size_t)n * 4 + 255) & ~(size_t)255) +
           (((size_t)nh * 4 + 255) & ~(size_t)255) + dt_scan_tmp_bytes(nh) + 256;
}

int dt_radix_sort_u32(dt_ctx *ctx, const uint32_t *kin, uint32_t *kout, const int32_t *vin, int32_t *vout, int64_t n,
                      int nbits, void *tmp) {
    return radix_sort_pairs<uint32_t>(ctx, kin, kout, vin, vout, n, nbits, tmp);
}

int dt_radix_sort_u64(dt_ctx *ctx, const uint64_t *kin, uint64_t *kout, const int32_t *vin, int32_t *vout, int64_t n,
                      int nbits, void *tmp) {
    return radix_sort_pairs<uint64_t>(ctx, kin, kout, vin, vout, n, nbits, tmp);
}
