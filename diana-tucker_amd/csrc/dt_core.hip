// Context, device memory, element conversion and timing of libdiana_hip.so.
// Replaces the reference's per-tensor Operator allocation
// (template/operator/operator_cpu.tpp:6-17) and the Summary timers
// (src/summary.cpp:29-69) for a single rank.
#include "dt_internal.hpp"

#include <algorithm>

int dt_fail(dt_ctx *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

void *dt_ws(dt_ctx *ctx, int slot, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (ctx->ws_size[slot] >= bytes) return ctx->ws[slot];
    if (ctx->ws[slot]) {
        hipStreamSynchronize(ctx->stream);
        hipFree(ctx->ws[slot]);
        ctx->ws[slot] = nullptr;
        ctx->ws_size[slot] = 0;
    }
    size_t sz = std::max(bytes, (size_t)(ctx->ws_size[slot] * 5 / 4));
    sz = (sz + 255) & ~(size_t)255;
    void *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        dt_fail(ctx, DT_ENOMEM, "workspace slot %d: hipMalloc(%zu) failed", slot, sz);
        return nullptr;
    }
    ctx->ws[slot] = p;
    ctx->ws_size[slot] = sz;
    return p;
}

hipEvent_t dt_get_event(dt_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

void dt_collect_timing(dt_ctx *ctx) {
    for (auto &p : ctx->pending) {
        hipEventSynchronize(p.b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, p.a, p.b);
        auto &s = ctx->stats[p.name];
        s.count += 1;
        s.ms += ms;
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.clear();
}

extern "C" {

int dt_create(dt_ctx **out, int device) {
    if (!out) return DT_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DT_EHIP;
    if (device < 0 || device >= n) return DT_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return DT_EHIP;
    dt_ctx *ctx = new dt_ctx();
    ctx->device = device;
    // Default-priority stream only.
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return DT_EHIP;
    }
    hipEventCreate(&ctx->t0);
    hipEventCreate(&ctx->t1);
    // side stream for overlapped work: default priority, like the main one;
    // without it side work runs on the main stream
    if (hipStreamCreateWithFlags(&ctx->side_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->side_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->side_done, hipEventDisableTiming) != hipSuccess) {
        if (ctx->side_stream) hipStreamDestroy(ctx->side_stream);
        ctx->side_stream = nullptr;
    }
    if (hipHostMalloc((void **)&ctx->pinned, 4096, hipHostMallocDefault) != hipSuccess) {
        ctx->pinned = nullptr;
    }
    // the allocation-free entry points' scratch address on this device (its
    // failure only disables them, with the reason kept for their error)
    if (dt_noalloc_setup(ctx) != DT_OK) {
        ctx->noalloc_err = ctx->err;
        ctx->err.clear();
    }
    *out = ctx;
    return DT_OK;
}

int dt_destroy(dt_ctx *ctx) {
    if (!ctx) return DT_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    dt_collect_timing(ctx);
    for (int i = 0; i < dt_ctx::kSlots; i++)
        if (ctx->ws[i]) hipFree(ctx->ws[i]);
    for (auto e : ctx->event_pool) hipEventDestroy(e);
    if (ctx->t0) hipEventDestroy(ctx->t0);
    if (ctx->t1) hipEventDestroy(ctx->t1);
    if (ctx->pinned) hipHostFree(ctx->pinned);
    if (ctx->side_stream) {
        hipStreamSynchronize(ctx->side_stream);
        hipStreamDestroy(ctx->side_stream);
    }
    if (ctx->side_fork) hipEventDestroy(ctx->side_fork);
    if (ctx->side_done) hipEventDestroy(ctx->side_done);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return DT_OK;
}

const char *dt_last_error(const dt_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int dt_malloc(dt_ctx *ctx, void **ptr, size_t bytes) {
    DT_CHECK_ARG(ctx && ptr, "dt_malloc: null argument");
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return dt_fail(ctx, DT_ENOMEM, "dt_malloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return DT_OK;
}

int dt_free(dt_ctx *ctx, void *ptr) {
    if (!ptr) return DT_OK;
    DT_HIP(hipStreamSynchronize(ctx->stream));
    DT_HIP(hipFree(ptr));
    return DT_OK;
}

int dt_h2d(dt_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return DT_OK;
    DT_CHECK_ARG(dst && src, "dt_h2d: null pointer");
    DT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    return DT_OK;
}

int dt_d2h(dt_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return DT_OK;
    DT_CHECK_ARG(dst && src, "dt_d2h: null pointer");
    DT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    return DT_OK;
}

int dt_d2d(dt_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return DT_OK;
    DT_CHECK_ARG(dst && src, "dt_d2d: null pointer");
    DT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return DT_OK;
}

int dt_sync(dt_ctx *ctx) {
    DT_HIP(hipStreamSynchronize(ctx->stream));
    return DT_OK;
}

int dt_timing_enable(dt_ctx *ctx, int on) {
    ctx->timing = on != 0;
    return DT_OK;
}

int dt_timing_filter(dt_ctx *ctx, const char *prefix) {
    ctx->timing_prefixes.clear();
    const std::string all = prefix ? prefix : "";
    size_t a = 0;
    while (a < all.size()) {
        size_t b = all.find(',', a);
        if (b == std::string::npos) b = all.size();
        if (b > a) ctx->timing_prefixes.push_back(all.substr(a, b - a));
        a = b + 1;
    }
    return DT_OK;
}

int dt_timing_reset(dt_ctx *ctx) {
    DT_HIP(hipStreamSynchronize(ctx->stream));
    dt_collect_timing(ctx);
    ctx->stats.clear();
    return DT_OK;
}

int dt_timing_get(dt_ctx *ctx, const char *kernel, int64_t *count, double *total_ms) {
    DT_HIP(hipStreamSynchronize(ctx->stream));
    dt_collect_timing(ctx);
    auto it = ctx->stats.find(kernel ? kernel : "");
    if (count) *count = it == ctx->stats.end() ? 0 : it->second.count;
    if (total_ms) *total_ms = it == ctx->stats.end() ? 0.0 : it->second.ms;
    return DT_OK;
}

int dt_timing_list(dt_ctx *ctx, char *buf, size_t buflen) {
    DT_HIP(hipStreamSynchronize(ctx->stream));
    dt_collect_timing(ctx);
    std::string s;
    char line[256];
    for (auto &kv : ctx->stats) {
        snprintf(line, sizeof line, "%s %lld %.6f\n", kv.first.c_str(), (long long)kv.second.count,
                 kv.second.ms);
        s += line;
    }
    if (buf && buflen) {
        size_t n = std::min(buflen - 1, s.size());
        memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return DT_OK;
}

int dt_timer_start(dt_ctx *ctx) {
    DT_HIP(hipEventRecord(ctx->t0, ctx->stream));
    return DT_OK;
}

int dt_timer_stop(dt_ctx *ctx, double *elapsed_ms) {
    DT_HIP(hipEventRecord(ctx->t1, ctx->stream));
    DT_HIP(hipEventSynchronize(ctx->t1));
    float ms = 0.f;
    DT_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    if (elapsed_ms) *elapsed_ms = ms;
    return DT_OK;
}

}  // extern "C"

// ---- elementwise kernels -----------------------------------------------------
template <typename T>
__device__ inline double load_as_f64(const T *p, size_t i);
template <>
__device__ inline double load_as_f64<double>(const double *p, size_t i) { return p[i]; }
template <>
__device__ inline double load_as_f64<float>(const float *p, size_t i) { return (double)p[i]; }
template <>
__device__ inline double load_as_f64<uint16_t>(const uint16_t *p, size_t i) {
    return (double)bf16_to_f32(p[i]);
}

template <typename T>
__device__ inline void store_from_f64(T *p, size_t i, double v);
template <>
__device__ inline void store_from_f64<double>(double *p, size_t i, double v) { p[i] = v; }
template <>
__device__ inline void store_from_f64<float>(float *p, size_t i, double v) { p[i] = (float)v; }
template <>
__device__ inline void store_from_f64<uint16_t>(uint16_t *p, size_t i, double v) {
    p[i] = f32_to_bf16((float)v);
}

template <typename TD, typename TS>
__global__ void convert_kernel(TD *__restrict__ dst, const TS *__restrict__ src, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        store_from_f64<TD>(dst, i, load_as_f64<TS>(src, i));
}

__global__ void zero_kernel(uint32_t *__restrict__ p, size_t nwords) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) p[i] = 0u;
}

__global__ void zero_bytes_kernel(uint8_t *__restrict__ p, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

static inline dim3 grid_for(size_t n, int block = 256) {
    size_t g = (n + block - 1) / block;
    if (g > 65536 * 4) g = 65536 * 4;
    if (g == 0) g = 1;
    return dim3((unsigned)g);
}

template <typename TD, typename TS>
static int launch_convert(dt_ctx *ctx, void *dst, const void *src, size_t n) {
    DT_LAUNCH("convert", (convert_kernel<TD, TS>), grid_for(n), dim3(256), 0, (TD *)dst,
              (const TS *)src, n);
    return DT_OK;
}

extern "C" int dt_convert(dt_ctx *ctx, void *dst, int dt_dst, const void *src, int dt_src,
                          size_t n) {
    if (n == 0) return DT_OK;
    DT_CHECK_ARG(dst && src, "dt_convert: null pointer");
#define DT_CONV_CASE(a, b, TA, TB) \
    if (dt_dst == a && dt_src == b) return launch_convert<TA, TB>(ctx, dst, src, n);
    DT_CONV_CASE(DT_F64, DT_F64, double, double)
    DT_CONV_CASE(DT_F64, DT_F32, double, float)
    DT_CONV_CASE(DT_F64, DT_BF16, double, uint16_t)
    DT_CONV_CASE(DT_F32, DT_F64, float, double)
    DT_CONV_CASE(DT_F32, DT_F32, float, float)
    DT_CONV_CASE(DT_F32, DT_BF16, float, uint16_t)
    DT_CONV_CASE(DT_BF16, DT_F64, uint16_t, double)
    DT_CONV_CASE(DT_BF16, DT_F32, uint16_t, float)
    DT_CONV_CASE(DT_BF16, DT_BF16, uint16_t, uint16_t)
#undef DT_CONV_CASE
    return dt_fail(ctx, DT_EUNSUPPORTED, "dt_convert: unsupported types %d <- %d", dt_dst, dt_src);
}

extern "C" int dt_zero(dt_ctx *ctx, void *dev, size_t bytes) {
    if (bytes == 0) return DT_OK;
    DT_CHECK_ARG(dev, "dt_zero: null pointer");
    // A fill kernel rather than hipMemsetAsync, so callers may capture it in a graph.
    if ((bytes & 3) == 0 && (((uintptr_t)dev) & 3) == 0) {
        DT_LAUNCH("zero", zero_kernel, grid_for(bytes / 4), dim3(256), 0, (uint32_t *)dev, bytes / 4);
    } else {
        DT_LAUNCH("zero", zero_bytes_kernel, grid_for(bytes), dim3(256), 0, (uint8_t *)dev, bytes);
    }
    return DT_OK;
}

int dt_copy_f64(dt_ctx *ctx, double *dst, const double *src, size_t n) {
    if (n == 0 || dst == src) return DT_OK;
    DT_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return DT_OK;
}
