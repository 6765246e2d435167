// Gather bandwidth of random 64-byte factor rows (the sparse TTMc's operand
// reads): rows of 16 FP32 values drawn at random from a table of 16 KB
// (L1-resident), 1 MB (L2-resident) or 64 MB (beyond the L2s), read by one
// wave-wide load per 16 rows (four lanes per row, 16 bytes each: the TTMc's
// layout), per 4 rows (16 lanes per row, 4 bytes each), per 8 rows (8 lanes,
// 8 bytes), or one row piece per lane, next to a coalesced stream of the same
// table.  Partially used 16-row loads show what a load's unused lanes cost:
// 4 of its 16 rows read with the other lanes masked off (a divergent branch),
// with the other lanes out of a buffer resource's range (they return zero and
// touch no memory), or with the 16 row slots holding 4 distinct rows; a load
// with every lane out of range gives the floor of one load instruction.
// Eight workgroups of four waves per CU, 200 x 8 loads per wave; the
// row numbers come from a hash of (wave, iteration, load, lane group) masked
// to the table, so every access stays inside it.  Prints, per case, the time
// per wave-wide load instruction per CU (2.4 GHz clock) and the row bytes
// delivered per clock per CU.
//
//     make -C diana-tucker_amd probes && diana-tucker_amd/probes/gather_probe
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));

namespace {

constexpr unsigned kMaxRows = 1u << 20;  // 64 MB of 64-byte rows
constexpr int kIters = 200, kLoads = 8;

__device__ __forceinline__ unsigned mix(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// four lanes per row, 16 bytes each: 16 rows per load
__global__ __launch_bounds__(256) void rows_x4(const float *__restrict__ T, unsigned mask, float *out) {
    const int lane = threadIdx.x & 63;
    const unsigned w = blockIdx.x * 4 + (threadIdx.x >> 6);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int it = 0; it < kIters; it++)
#pragma unroll
        for (int s = 0; s < kLoads; s++) {
            const unsigned r = mix(w * 977 + it * 131 + s * 17 + (lane >> 2)) & mask;
            acc += *(const f4 *)(T + r * 16 + 4 * (lane & 3));
        }
    out[blockIdx.x * 256 + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

// 16-row loads of which only 4 rows are used.  MODE 0: the other lanes
// masked off; MODE 1: the other lanes out of a buffer resource's range;
// MODE 2: every lane out of range; MODE 3: four distinct rows, each read by
// four lane groups.
template <int MODE>
__global__ __launch_bounds__(256) void rows_x4_part(const float *__restrict__ T, unsigned mask, float *out) {
    const int lane = threadIdx.x & 63, p4 = lane >> 2, c4 = lane & 3;
    const unsigned w = blockIdx.x * 4 + (threadIdx.x >> 6);
    // byte range of the whole table; offsets at or past it read zero
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)T, (short)0, (int)(kMaxRows * 64), 0x00020000);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int it = 0; it < kIters; it++)
#pragma unroll
        for (int s = 0; s < kLoads; s++) {
            const unsigned r = mix(w * 977 + it * 131 + s * 17 + (MODE == 3 ? (p4 & 3) : p4)) & mask;
            if (MODE == 0) {
                if (p4 < 4) acc += *(const f4 *)(T + r * 16 + 4 * c4);
            } else if (MODE == 1 || MODE == 2) {
                // out-of-range offsets still depend on r, so that no two loads are the same
                const unsigned off = (MODE == 1 && p4 < 4) ? r * 64 + 16 * c4 : 0x80000000u + r * 64 + 16 * c4;
                acc += __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
            } else {
                acc += *(const f4 *)(T + r * 16 + 4 * c4);
            }
        }
    out[blockIdx.x * 256 + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

// eight lanes per row, 8 bytes each: 8 rows per load
__global__ __launch_bounds__(256) void rows_x2(const float *__restrict__ T, unsigned mask, float *out) {
    const int lane = threadIdx.x & 63;
    const unsigned w = blockIdx.x * 4 + (threadIdx.x >> 6);
    f2 acc = {0.f, 0.f};
    for (int it = 0; it < kIters; it++)
#pragma unroll
        for (int s = 0; s < kLoads; s++) {
            const unsigned r = mix(w * 977 + it * 131 + s * 17 + (lane >> 3)) & mask;
            acc += *(const f2 *)(T + r * 16 + 2 * (lane & 7));
        }
    out[blockIdx.x * 256 + threadIdx.x] = acc.x + acc.y;
}

// sixteen lanes per row, 4 bytes each: 4 rows per load
__global__ __launch_bounds__(256) void rows_x1(const float *__restrict__ T, unsigned mask, float *out) {
    const int lane = threadIdx.x & 63;
    const unsigned w = blockIdx.x * 4 + (threadIdx.x >> 6);
    float acc = 0.f;
    for (int it = 0; it < kIters; it++)
#pragma unroll
        for (int s = 0; s < kLoads; s++) {
            const unsigned r = mix(w * 977 + it * 131 + s * 17 + (lane >> 4)) & mask;
            acc += T[r * 16 + (lane & 15)];
        }
    out[blockIdx.x * 256 + threadIdx.x] = acc;
}

// one lane per row, a 16-byte quarter per load: 64 rows per load
__global__ __launch_bounds__(256) void rows_lane(const float *__restrict__ T, unsigned mask, float *out) {
    const int lane = threadIdx.x & 63;
    const unsigned w = blockIdx.x * 4 + (threadIdx.x >> 6);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int it = 0; it < kIters; it++)
#pragma unroll
        for (int s = 0; s < kLoads; s++) {
            const unsigned r = mix(w * 977 + it * 131 + (s >> 2) * 17 + lane) & mask;
            acc += *(const f4 *)(T + r * 16 + 4 * (s & 3));
        }
    out[blockIdx.x * 256 + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

// coalesced: 16 consecutive rows (1 KB) per load at a random 1 KB-aligned place
__global__ __launch_bounds__(256) void stream_x4(const float *__restrict__ T, unsigned mask, float *out) {
    const int lane = threadIdx.x & 63;
    const unsigned w = blockIdx.x * 4 + (threadIdx.x >> 6);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int it = 0; it < kIters; it++)
#pragma unroll
        for (int s = 0; s < kLoads; s++) {
            const unsigned r = (mix(w * 977 + it * 131 + s * 17) & (mask >> 4)) * 16 + (lane >> 2);
            acc += *(const f4 *)(T + r * 16 + 4 * (lane & 3));
        }
    out[blockIdx.x * 256 + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

#define CHECK(x)                                                                                      \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) {                                                                       \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                              \
            std::exit(1);                                                                             \
        }                                                                                             \
    } while (0)

}  // namespace

int main() {
    int cus = 0;
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    float *T = nullptr, *out = nullptr;
    CHECK(hipMalloc(&T, (size_t)kMaxRows * 64));
    CHECK(hipMemset(T, 0, (size_t)kMaxRows * 64));
    const int wg = cus * 8;
    CHECK(hipMalloc(&out, (size_t)wg * 256 * 4));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    struct Kern {
        const char *name;
        void (*f)(const float *, unsigned, float *);
        double row_bytes_per_load;
    } ks[] = {{"4 lanes/row x 16 B", rows_x4, 1024}, {"8 lanes/row x 8 B", rows_x2, 512},
              {"16 lanes/row x 4 B", rows_x1, 256}, {"1 lane/row x 16 B", rows_lane, 1024},
              {"coalesced 1 KB", stream_x4, 1024},
              {"4 of 16 rows, masked", rows_x4_part<0>, 256}, {"4 of 16 rows, 12 OOB", rows_x4_part<1>, 256},
              {"all lanes OOB", rows_x4_part<2>, 0}, {"4 rows x 4 slots", rows_x4_part<3>, 256}};
    struct Table {
        const char *name;
        unsigned rows;
    } ts[] = {{"16 KB table", 256}, {"1 MB table", 1u << 14}, {"64 MB table", kMaxRows}};
    std::printf("%d CUs, %d workgroups of 4 waves, %d loads per wave\n", cus, wg, kIters * kLoads);
    for (const Table &t : ts) {
        std::printf("-- %s\n", t.name);
        for (const Kern &k : ks) {
            float ms = 0.f;
            for (int rep = 0; rep < 2; rep++) {  // the first run warms the caches
                CHECK(hipEventRecord(a));
                hipLaunchKernelGGL(k.f, dim3(wg), dim3(256), 0, 0, T, t.rows - 1, out);
                CHECK(hipGetLastError());
                CHECK(hipEventRecord(b));
                CHECK(hipEventSynchronize(b));
                CHECK(hipEventElapsedTime(&ms, a, b));
            }
            const double loads_per_cu = (double)wg * 4 * kIters * kLoads / cus;
            const double clk = ms * 1e-3 * 2.4e9;
            std::printf("  %-20s %7.3f ms  %6.1f clk per load per CU  %5.1f B/clk/CU  %6.2f TB/s\n", k.name, ms,
                        clk / loads_per_cu, loads_per_cu * k.row_bytes_per_load / clk,
                        loads_per_cu * cus * k.row_bytes_per_load / (ms * 1e-3) / 1e12);
        }
    }
    CHECK(hipFree(T));
    CHECK(hipFree(out));
    return 0;
}
