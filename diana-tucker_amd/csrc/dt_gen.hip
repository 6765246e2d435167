// Seeded generators: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) with
// Box-Muller normals, and the synthetic Tucker model used by the benchmarks.
//
// Element i of stream s under seed k is a pure function of (k, s, i):
//   call q = i / 2, counter = (q_lo, q_hi, s_lo, s_hi), key = (k_lo, k_hi)
//   (x0,x1,x2,x3) = philox4x32_10(counter, key)
//   a = x0 | x1 << 32,  b = x2 | x3 << 32
//   u1 = ((a >> 11) + 1) 2^-53 in (0, 1],  u2 = (b >> 11) 2^-53 in [0, 1)
//   normal(i) = sqrt(-2 ln u1) * (i even ? cos(2 pi u2) : sin(2 pi u2))
//   uniform(i) = (((i even ? a : b) >> 11) + 1) 2^-53 in (0, 1]
// The oracle (oracle/philox.py) restates the same integer stream.
//
// The reference has only glibc rand() + the polar method (src/util.cpp:10-18,
// operator_double.cpp:66-72); that generator is reproduced on the host side
// (diana-tucker_amd/host) for reference-style random starts.
#include "dt_internal.hpp"

#include <cmath>

#include "dt_philox.hpp"

template <typename T>
__device__ inline void store_val(T *p, int64_t i, double v);
template <>
__device__ inline void store_val<double>(double *p, int64_t i, double v) { p[i] = v; }
template <>
__device__ inline void store_val<float>(float *p, int64_t i, double v) { p[i] = (float)v; }
template <>
__device__ inline void store_val<uint16_t>(uint16_t *p, int64_t i, double v) {
    p[i] = f32_to_bf16((float)v);
}

template <typename T>
__global__ void gen_normal_kernel(T *__restrict__ out, int64_t n, uint64_t seed, uint64_t stream) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        store_val<T>(out, i, normal_at(seed, stream, (uint64_t)i));
}

template <typename T>
__global__ void gen_uniform_kernel(T *__restrict__ out, int64_t n, uint64_t seed, uint64_t stream) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        store_val<T>(out, i, uniform_at(seed, stream, (uint64_t)i));
}

static inline dim3 gs_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 262144) g = 262144;
    if (g < 1) g = 1;
    return dim3((unsigned)g);
}

int dt_fill_normal_f64(dt_ctx *ctx, double *out, size_t n, uint64_t seed, uint64_t stream) {
    DT_LAUNCH("gen_normal", gen_normal_kernel<double>, gs_grid((int64_t)n), dim3(256), 0, out,
              (int64_t)n, seed, stream);
    return DT_OK;
}

extern "C" int dt_gen_normal(dt_ctx *ctx, void *out, int dt, size_t n, uint64_t seed, uint64_t stream) {
    DT_CHECK_ARG(ctx && (out || n == 0), "dt_gen_normal: null argument");
    if (n == 0) return DT_OK;
    if (dt == DT_F64) {
        DT_LAUNCH("gen_normal", gen_normal_kernel<double>, gs_grid((int64_t)n), dim3(256), 0,
                  (double *)out, (int64_t)n, seed, stream);
    } else if (dt == DT_F32) {
        DT_LAUNCH("gen_normal", gen_normal_kernel<float>, gs_grid((int64_t)n), dim3(256), 0,
                  (float *)out, (int64_t)n, seed, stream);
    } else if (dt == DT_BF16) {
        DT_LAUNCH("gen_normal", gen_normal_kernel<uint16_t>, gs_grid((int64_t)n), dim3(256), 0,
                  (uint16_t *)out, (int64_t)n, seed, stream);
    } else {
        return dt_fail(ctx, DT_EUNSUPPORTED, "dt_gen_normal: type %d", dt);
    }
    return DT_OK;
}

extern "C" int dt_gen_uniform(dt_ctx *ctx, void *out, int dt, size_t n, uint64_t seed,
                              uint64_t stream) {
    DT_CHECK_ARG(ctx && (out || n == 0), "dt_gen_uniform: null argument");
    if (n == 0) return DT_OK;
    if (dt == DT_F64) {
        DT_LAUNCH("gen_uniform", gen_uniform_kernel<double>, gs_grid((int64_t)n), dim3(256), 0,
                  (double *)out, (int64_t)n, seed, stream);
    } else if (dt == DT_F32) {
        DT_LAUNCH("gen_uniform", gen_uniform_kernel<float>, gs_grid((int64_t)n), dim3(256), 0,
                  (float *)out, (int64_t)n, seed, stream);
    } else {
        return dt_fail(ctx, DT_EUNSUPPORTED, "dt_gen_uniform: type %d", dt);
    }
    return DT_OK;
}

// ---- Tucker model --------------------------------------------------------------
// Core decay: G[r_0, ..., r_{N-1}] *= prod_n decay^{r_n}
__global__ void core_decay_kernel(double *__restrict__ G, int ndim, const int64_t *__restrict__ R,
                                  int64_t total, double decay) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        int64_t t = idx;
        int64_t e = 0;
        for (int d = 0; d < ndim; d++) {
            e += t % R[d];
            t /= R[d];
        }
        G[idx] *= pow(decay, (double)e);
    }
}

// Y[f, j, b] = sum_k T[f, k, b] U[j + J k]   (FP64 accumulate)
//   U here is J x K col-major (the true factor I_n x R_n: J = I_n, K = R_n).
template <typename TY>
__global__ __launch_bounds__(256) void expand_kernel(const double *__restrict__ T,
                                                     const double *__restrict__ U,
                                                     TY *__restrict__ Y, int64_t F, int64_t K,
                                                     int64_t B, int64_t J) {
    const int64_t total = F * J * B;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const int64_t f = idx % F;
        const int64_t t = idx / F;
        const int64_t j = t % J;
        const int64_t b = t / J;
        const double *tp = T + f + F * K * b;
        double acc = 0.0;
        for (int64_t k = 0; k < K; k++) acc += tp[F * k] * U[j + J * k];
        store_val<TY>(Y, idx, acc);
    }
}

__device__ inline double block_sum_d(double v) {
    __shared__ double red[16];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)((blockDim.x + 63) >> 6); i++) s += red[i];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(256) void noise_sumsq_kernel(int64_t n, uint64_t seed, uint64_t stream,
                                                          double *__restrict__ part) {
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double e = normal_at(seed, stream, (uint64_t)i);
        s += e * e;
    }
    s = block_sum_d(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void sum_part_kernel(const double *__restrict__ part, int n, double *__restrict__ out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    s = block_sum_d(s);
    if (threadIdx.x == 0) out[0] = s;
}

__global__ __launch_bounds__(256) void noise_add_kernel(float *__restrict__ A,
                                                        uint16_t *__restrict__ Ab, int64_t n,
                                                        uint64_t seed, uint64_t stream,
                                                        double scale) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = (double)A[i] + scale * normal_at(seed, stream, (uint64_t)i);
        const float fv = (float)v;
        A[i] = fv;
        if (Ab) Ab[i] = f32_to_bf16(fv);
    }
}

extern "C" int dt_gen_tucker(dt_ctx *ctx, const dt_tucker_spec *spec, float *A_f32, uint16_t *A_bf16,
                             double *const *U_true, double *normX_host, double *normA_host) {
    DT_CHECK_ARG(ctx && spec && A_f32, "dt_gen_tucker: null argument (A_f32 is required)");
    const int N = spec->ndim;
    DT_CHECK_ARG(N >= 2 && N <= DT_MAX_NDIM, "dt_gen_tucker: ndim %d", N);
    int64_t I[DT_MAX_NDIM], R[DT_MAX_NDIM];
    int64_t total = 1, coresz = 1;
    for (int n = 0; n < N; n++) {
        I[n] = (int64_t)spec->shape[n];
        R[n] = (int64_t)spec->ranks[n];
        DT_CHECK_ARG(R[n] >= 1 && R[n] <= I[n], "dt_gen_tucker: rank %lld > size %lld",
                     (long long)R[n], (long long)I[n]);
        total *= I[n];
        coresz *= R[n];
    }
    const uint64_t seed = spec->seed;
    // factors
    std::vector<double *> U(N, nullptr);
    std::vector<void *> owned;
    auto cleanup = [&]() {
        hipStreamSynchronize(ctx->stream);
        for (void *p : owned) hipFree(p);
    };
    for (int n = 0; n < N; n++) {
        double *p = nullptr;
        if (hipMalloc(&p, (size_t)(I[n] * R[n]) * sizeof(double)) != hipSuccess) {
            cleanup();
            return dt_fail(ctx, DT_ENOMEM, "dt_gen_tucker: factor allocation");
        }
        owned.push_back(p);
        U[n] = p;
        int st = dt_fill_normal_f64(ctx, p, (size_t)(I[n] * R[n]), seed, 2000 + (uint64_t)n);
        if (st == DT_OK) st = dt_qr_f64(ctx, p, nullptr, p, (size_t)I[n], (size_t)R[n]);
        if (st != DT_OK) {
            cleanup();
            return st;
        }
    }
    // core
    double *G = nullptr;
    int64_t *dR = nullptr;
    if (hipMalloc(&G, (size_t)coresz * sizeof(double)) != hipSuccess ||
        hipMalloc(&dR, sizeof(int64_t) * DT_MAX_NDIM) != hipSuccess) {
        cleanup();
        return dt_fail(ctx, DT_ENOMEM, "dt_gen_tucker: core allocation");
    }
    owned.push_back(G);
    owned.push_back(dR);
    int st = dt_fill_normal_f64(ctx, G, (size_t)coresz, seed, 1000);
    if (st != DT_OK) {
        cleanup();
        return st;
    }
    if (spec->decay > 0.0) {
        hipMemcpyAsync(dR, R, sizeof(int64_t) * N, hipMemcpyHostToDevice, ctx->stream);
        hipLaunchKernelGGL(core_decay_kernel, gs_grid(coresz), dim3(256), 0, ctx->stream, G, N, dR,
                           coresz, spec->decay);
    }
    // expanding chain: modes 0..N-2 in FP64, last mode straight into A_f32
    int64_t cur_shape[DT_MAX_NDIM];
    for (int n = 0; n < N; n++) cur_shape[n] = R[n];
    const double *src = G;
    for (int n = 0; n < N; n++) {
        int64_t F = 1, B = 1;
        for (int i = 0; i < n; i++) F *= cur_shape[i];
        for (int i = n + 1; i < N; i++) B *= cur_shape[i];
        const int64_t K = R[n], J = I[n];
        if (n == N - 1) {
            hipLaunchKernelGGL(expand_kernel<float>, gs_grid(F * J * B), dim3(256), 0, ctx->stream, src,
                               U[n], A_f32, F, K, B, J);
        } else {
            double *dst = nullptr;
            if (hipMalloc(&dst, (size_t)(F * J * B) * sizeof(double)) != hipSuccess) {
                cleanup();
                return dt_fail(ctx, DT_ENOMEM, "dt_gen_tucker: chain allocation");
            }
            owned.push_back(dst);
            hipLaunchKernelGGL(expand_kernel<double>, gs_grid(F * J * B), dim3(256), 0, ctx->stream,
                               src, U[n], dst, F, K, B, J);
            src = dst;
        }
        cur_shape[n] = I[n];
    }
    if (hipGetLastError() != hipSuccess) {
        cleanup();
        return dt_fail(ctx, DT_EHIP, "dt_gen_tucker: kernel launch failed");
    }
    // norms and noise
    double normX2 = 0.0;
    st = dt_sumsq_dev(ctx, A_f32, DT_F32, (size_t)total, &normX2);
    if (st != DT_OK) {
        cleanup();
        return st;
    }
    const int nb = 2048;
    double *part = (double *)dt_ws(ctx, WS_GEN, (nb + 8) * sizeof(double));
    if (!part) {
        cleanup();
        return DT_ENOMEM;
    }
    double scale = 0.0;
    if (spec->noise > 0.0) {
        hipLaunchKernelGGL(noise_sumsq_kernel, dim3(nb), dim3(256), 0, ctx->stream, total, seed,
                           (uint64_t)3000, part);
        hipLaunchKernelGGL(sum_part_kernel, dim3(1), dim3(256), 0, ctx->stream, part, nb, part + nb);
        hipMemcpyAsync(ctx->pinned, part + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        hipStreamSynchronize(ctx->stream);
        const double normE2 = ctx->pinned[0];
        scale = spec->noise * std::sqrt(normX2 / normE2);
    }
    hipLaunchKernelGGL(noise_add_kernel, gs_grid(total), dim3(256), 0, ctx->stream, A_f32, A_bf16,
                       total, seed, (uint64_t)3000, scale);
    if (hipGetLastError() != hipSuccess) {
        cleanup();
        return dt_fail(ctx, DT_EHIP, "dt_gen_tucker: noise kernel launch failed");
    }
    double normA2 = 0.0;
    st = dt_sumsq_dev(ctx, A_f32, DT_F32, (size_t)total, &normA2);
    if (st != DT_OK) {
        cleanup();
        return st;
    }
    if (U_true) {
        for (int n = 0; n < N; n++)
            if (U_true[n])
                hipMemcpyAsync(U_true[n], U[n], (size_t)(I[n] * R[n]) * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream);
    }
    if (normX_host) *normX_host = std::sqrt(normX2);
    if (normA_host) *normA_host = std::sqrt(normA2);
    cleanup();
    return DT_OK;
}
