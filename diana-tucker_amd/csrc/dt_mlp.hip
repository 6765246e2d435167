// Row-wise MLP of the learned start: forward, backward and Adam on the GPU.
//
// The reference has no learned component; this is the learned start that
// replaces the HOSVD initialisation.  One network maps each row x (d_in features of a mode's
// sketch) to d_out values:
//     h1 = relu(x W1 + b1),  h2 = relu(h1 W2 + b2),  z = x[:d_out] + h2 W3 + b3
// W1: d_in x H, W2: H x H, W3: H x d_out, all row-major; the residual term
// makes a zero-initialised last layer reproduce the sketch's own leading
// directions.  Weights are kept in FP32 (Adam master copy) and rounded to
// bf16 wherever they enter a product; every product runs on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation (bf16 weights, FP32
// accumulation).  Activations are stored in FP32 and rounded to bf16 as MFMA
// operands.  Loss: sum over rows of |z - t|^2 divided by the number of rows.
#include "dt_internal.hpp"

#include <cmath>
#include <vector>

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int MT = 64;    // output tile (rows and columns)
constexpr int KT = 32;    // k step
constexpr int KP = 40;    // LDS row stride in bf16 (80 bytes)

// C (M x N, row-major, ldc) = op(A) op(B) [+ bias] [relu] [* (mask > 0)] [+ res]
//   TA: op(A)[m][k] = A[k * lda + m]   else A[m * lda + k]
//   TB: op(B)[k][n] = B[n * ldb + k]   else B[k * ldb + n]
struct Epi {
    const float *bias;  // N, or null
    int relu;
    const float *mask;  // M x N (ldm), multiply by (mask > 0), or null
    int64_t ldm;
    const float *res;   // M x (>= N) (ldr), added, or null
    int64_t ldr;
};

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void mlp_gemm(const float *__restrict__ A, int64_t lda,
                                                const float *__restrict__ B, int64_t ldb,
                                                float *__restrict__ C, int64_t ldc, int64_t M, int64_t N,
                                                int64_t K, Epi epi) {
    __shared__ uint16_t As[MT][KP];
    __shared__ uint16_t Bs[MT][KP];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int64_t m0 = (int64_t)blockIdx.y * MT, n0 = (int64_t)blockIdx.x * MT;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int64_t k0 = 0; k0 < K; k0 += KT) {
#pragma unroll
        for (int q = 0; q < (MT * KT) / 256; q++) {
            const int e = tid + 256 * q;
            int m, kk;
            if (TA) {
                kk = e / MT;
                m = e % MT;
            } else {
                m = e / KT;
                kk = e % KT;
            }
            const int64_t gm = m0 + m, gk = k0 + kk;
            float v = 0.f;
            if (gm < M && gk < K) v = TA ? A[gk * lda + gm] : A[gm * lda + gk];
            As[m][kk] = f32_to_bf16(v);
        }
#pragma unroll
        for (int q = 0; q < (MT * KT) / 256; q++) {
            const int e = tid + 256 * q;
            int n, kk;
            if (TB) {
                n = e / KT;
                kk = e % KT;
            } else {
                kk = e / MT;
                n = e % MT;
            }
            const int64_t gn = n0 + n, gk = k0 + kk;
            float v = 0.f;
            if (gn < N && gk < K) v = TB ? B[gn * ldb + gk] : B[gk * ldb + gn];
            Bs[n][kk] = f32_to_bf16(v);
        }
        __syncthreads();
        u32x4 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) a[i] = *(const u32x4 *)&As[wm + 16 * i + c][8 * g];
#pragma unroll
        for (int j = 0; j < 2; j++) b[j] = *(const u32x4 *)&Bs[wn + 16 * j + c][8 * g];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[i]),
                                                                    __builtin_bit_cast(bf16x8, b[j]),
                                                                    acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t m = m0 + wm + 16 * i + 4 * g + v;
                const int64_t n = n0 + wn + 16 * j + c;
                if (m < M && n < N) {
                    float x = acc[i][j][v];
                    if (epi.bias) x += epi.bias[n];
                    if (epi.relu) x = x > 0.f ? x : 0.f;
                    if (epi.mask && !(epi.mask[m * epi.ldm + n] > 0.f)) x = 0.f;
                    if (epi.res) x += epi.res[m * epi.ldr + n];
                    C[m * ldc + n] = x;
                }
            }
}

// part[chunk][n] = sum over the chunk's rows of D[m][n]  (chunks of 256 rows)
__global__ void colsum_partial(const float *__restrict__ D, int64_t M, int64_t N, float *__restrict__ part) {
    const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int64_t r0 = (int64_t)blockIdx.y * 256;
    const int64_t r1 = r0 + 256 < M ? r0 + 256 : M;
    float s = 0.f;
    for (int64_t m = r0; m < r1; m++) s += D[m * N + n];
    part[blockIdx.y * N + n] = s;
}
__global__ void colsum_final(const float *__restrict__ part, int nchunks, int64_t N, float *__restrict__ out) {
    const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int q = 0; q < nchunks; q++) s += part[q * N + n];
    out[n] = s;
}

// dZ = 2 (Z - T) / rows; partial sums of |Z - T|^2 per block (FP64)
__global__ void mse_grad(const float *__restrict__ Z, const float *__restrict__ T, int64_t n, float scale,
                         float *__restrict__ dZ, double *__restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float d = Z[i] - T[i];
        dZ[i] = 2.f * d * scale;
        s += (double)d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void adam_kernel(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ Mo,
                            float *__restrict__ Ve, int64_t n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float g = G[i];
        const float m = b1 * Mo[i] + (1.f - b1) * g;
        const float v = b2 * Ve[i] + (1.f - b2) * g * g;
        Mo[i] = m;
        Ve[i] = v;
        P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    }
}

inline dim3 ew_grid(int64_t n) { return dim3((unsigned)std::min<int64_t>((n + 255) / 256, 2048)); }

}  // namespace

struct dt_mlp {
    int din = 0, hid = 0, dout = 0;
    int64_t nparam = 0;
    // parameter offsets inside P / G / Mo / Ve
    int64_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Ve = nullptr;
    int64_t step = 0;
    // activations (rows x width), grown on demand
    int64_t cap_rows = 0;
    float *H1 = nullptr, *H2 = nullptr, *Z = nullptr, *D1 = nullptr, *D2 = nullptr, *DZ = nullptr;
    float *colpart = nullptr;
    double *lpart = nullptr;
};

static int mlp_gemm_launch(dt_ctx *ctx, bool ta, bool tb, const float *A, int64_t lda, const float *B,
                           int64_t ldb, float *C, int64_t ldc, int64_t M, int64_t N, int64_t K, Epi epi) {
    dim3 grid((unsigned)((N + MT - 1) / MT), (unsigned)((M + MT - 1) / MT));
    if (!ta && !tb) {
        DT_LAUNCH("mlp_gemm", (mlp_gemm<false, false>), grid, dim3(256), 0, A, lda, B, ldb, C, ldc, M, N, K, epi);
    } else if (!ta && tb) {
        DT_LAUNCH("mlp_gemm", (mlp_gemm<false, true>), grid, dim3(256), 0, A, lda, B, ldb, C, ldc, M, N, K, epi);
    } else if (ta && !tb) {
        DT_LAUNCH("mlp_gemm", (mlp_gemm<true, false>), grid, dim3(256), 0, A, lda, B, ldb, C, ldc, M, N, K, epi);
    } else {
        return dt_fail(ctx, DT_EUNSUPPORTED, "mlp_gemm: TT not used");
    }
    return DT_OK;
}

static int mlp_reserve(dt_ctx *ctx, dt_mlp *m, int64_t rows) {
    if (rows <= m->cap_rows) return DT_OK;
    DT_HIP(hipStreamSynchronize(ctx->stream));
    for (float **p : {&m->H1, &m->H2, &m->Z, &m->D1, &m->D2, &m->DZ, &m->colpart}) {
        if (*p) DT_HIP(hipFree(*p));
        *p = nullptr;
    }
    if (m->lpart) DT_HIP(hipFree(m->lpart));
    m->lpart = nullptr;
    const int64_t H = m->hid;
    const int64_t nchunk = (rows + 255) / 256;
    if (hipMalloc(&m->H1, rows * H * 4) != hipSuccess || hipMalloc(&m->H2, rows * H * 4) != hipSuccess ||
        hipMalloc(&m->Z, rows * m->dout * 4) != hipSuccess || hipMalloc(&m->D1, rows * H * 4) != hipSuccess ||
        hipMalloc(&m->D2, rows * H * 4) != hipSuccess || hipMalloc(&m->DZ, rows * m->dout * 4) != hipSuccess ||
        hipMalloc(&m->colpart, nchunk * std::max<int64_t>(H, m->dout) * 4) != hipSuccess ||
        hipMalloc(&m->lpart, 2048 * sizeof(double)) != hipSuccess)
        return dt_fail(ctx, DT_ENOMEM, "dt_mlp: cannot allocate activations for %lld rows", (long long)rows);
    m->cap_rows = rows;
    return DT_OK;
}

// Z (rows x d_out) of the current parameters; keeps H1/H2 for backward.
static int mlp_forward_internal(dt_ctx *ctx, dt_mlp *m, const float *X, int64_t rows, float *Z) {
    const int64_t H = m->hid;
    Epi e1{m->P + m->ob1, 1, nullptr, 0, nullptr, 0};
    DT_TRY(mlp_gemm_launch(ctx, false, false, X, m->din, m->P + m->oW1, H, m->H1, H, rows, H, m->din, e1));
    Epi e2{m->P + m->ob2, 1, nullptr, 0, nullptr, 0};
    DT_TRY(mlp_gemm_launch(ctx, false, false, m->H1, H, m->P + m->oW2, H, m->H2, H, rows, H, H, e2));
    Epi e3{m->P + m->ob3, 0, nullptr, 0, X, m->din};
    DT_TRY(mlp_gemm_launch(ctx, false, false, m->H2, H, m->P + m->oW3, m->dout, Z, m->dout, rows, m->dout, H, e3));
    return DT_OK;
}

static int colsum(dt_ctx *ctx, dt_mlp *m, const float *D, int64_t rows, int64_t N, float *out) {
    const int nchunk = (int)((rows + 255) / 256);
    dim3 g1((unsigned)((N + 255) / 256), (unsigned)nchunk);
    DT_LAUNCH("mlp_colsum", colsum_partial, g1, dim3(256), 0, D, rows, N, m->colpart);
    DT_LAUNCH("mlp_colsum", colsum_final, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, m->colpart, nchunk, N,
              out);
    return DT_OK;
}

extern "C" int dt_mlp_create(dt_ctx *ctx, dt_mlp **out, int d_in, int hidden, int d_out, uint64_t seed) {
    DT_CHECK_ARG(ctx && out, "dt_mlp_create: null argument");
    DT_CHECK_ARG(d_in >= 1 && hidden >= 1 && d_out >= 1 && d_out <= d_in, "dt_mlp_create: need 1 <= d_out <= d_in");
    dt_mlp *m = new dt_mlp();
    m->din = d_in;
    m->hid = hidden;
    m->dout = d_out;
    m->oW1 = 0;
    m->ob1 = m->oW1 + (int64_t)d_in * hidden;
    m->oW2 = m->ob1 + hidden;
    m->ob2 = m->oW2 + (int64_t)hidden * hidden;
    m->oW3 = m->ob2 + hidden;
    m->ob3 = m->oW3 + (int64_t)hidden * d_out;
    m->nparam = m->ob3 + d_out;
    const size_t bytes = (size_t)m->nparam * 4;
    if (hipMalloc(&m->P, bytes) != hipSuccess || hipMalloc(&m->G, bytes) != hipSuccess ||
        hipMalloc(&m->Mo, bytes) != hipSuccess || hipMalloc(&m->Ve, bytes) != hipSuccess) {
        dt_mlp_destroy(ctx, m);
        return dt_fail(ctx, DT_ENOMEM, "dt_mlp_create: out of memory");
    }
    // He-normal hidden layers (Philox stream 4000+layer), zero biases, zero last layer
    std::vector<float> h((size_t)m->nparam, 0.f);
    double *tmp = nullptr;
    const int64_t n1 = (int64_t)d_in * hidden, n2 = (int64_t)hidden * hidden;
    if (hipMalloc(&tmp, (size_t)std::max(n1, n2) * 8) != hipSuccess) {
        dt_mlp_destroy(ctx, m);
        return dt_fail(ctx, DT_ENOMEM, "dt_mlp_create: out of memory");
    }
    std::vector<double> t((size_t)std::max(n1, n2));
    int st = dt_fill_normal_f64(ctx, tmp, (size_t)n1, seed, 4000);
    if (st == DT_OK && hipMemcpyAsync(t.data(), tmp, n1 * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        st = DT_EHIP;
    if (st == DT_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) st = DT_EHIP;
    for (int64_t i = 0; st == DT_OK && i < n1; i++) h[(size_t)(m->oW1 + i)] = (float)(t[(size_t)i] * std::sqrt(2.0 / d_in));
    if (st == DT_OK) st = dt_fill_normal_f64(ctx, tmp, (size_t)n2, seed, 4001);
    if (st == DT_OK && hipMemcpyAsync(t.data(), tmp, n2 * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        st = DT_EHIP;
    if (st == DT_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) st = DT_EHIP;
    for (int64_t i = 0; st == DT_OK && i < n2; i++) h[(size_t)(m->oW2 + i)] = (float)(t[(size_t)i] * std::sqrt(2.0 / hidden));
    (void)hipFree(tmp);
    if (st == DT_OK && hipMemcpy(m->P, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) st = DT_EHIP;
    if (st == DT_OK && hipMemset(m->Mo, 0, bytes) != hipSuccess) st = DT_EHIP;
    if (st == DT_OK && hipMemset(m->Ve, 0, bytes) != hipSuccess) st = DT_EHIP;
    if (st == DT_OK && hipMemset(m->G, 0, bytes) != hipSuccess) st = DT_EHIP;
    if (st != DT_OK) {
        dt_mlp_destroy(ctx, m);
        return dt_fail(ctx, st, "dt_mlp_create: initialisation failed");
    }
    *out = m;
    return DT_OK;
}

extern "C" int dt_mlp_destroy(dt_ctx *ctx, dt_mlp *m) {
    (void)ctx;
    if (!m) return DT_OK;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    for (float *p : {m->P, m->G, m->Mo, m->Ve, m->H1, m->H2, m->Z, m->D1, m->D2, m->DZ, m->colpart})
        if (p) (void)hipFree(p);
    if (m->lpart) (void)hipFree(m->lpart);
    delete m;
    return DT_OK;
}

extern "C" int dt_mlp_shape(const dt_mlp *m, int *d_in, int *hidden, int *d_out, int64_t *nparam) {
    if (!m) return DT_EINVAL;
    if (d_in) *d_in = m->din;
    if (hidden) *hidden = m->hid;
    if (d_out) *d_out = m->dout;
    if (nparam) *nparam = m->nparam;
    return DT_OK;
}

extern "C" int dt_mlp_get_params(dt_ctx *ctx, const dt_mlp *m, float *host, int64_t n) {
    DT_CHECK_ARG(ctx && m && host && n == m->nparam, "dt_mlp_get_params: need %lld floats", m ? (long long)m->nparam : 0LL);
    DT_HIP(hipMemcpyAsync(host, m->P, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    return DT_OK;
}

extern "C" int dt_mlp_set_params(dt_ctx *ctx, dt_mlp *m, const float *host, int64_t n) {
    DT_CHECK_ARG(ctx && m && host && n == m->nparam, "dt_mlp_set_params: need %lld floats", m ? (long long)m->nparam : 0LL);
    DT_HIP(hipMemcpyAsync(m->P, host, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    DT_HIP(hipStreamSynchronize(ctx->stream));
    m->step = 0;
    DT_TRY(dt_zero(ctx, m->Mo, (size_t)n * 4));
    return dt_zero(ctx, m->Ve, (size_t)n * 4);
}

int dt_mlp_forward_dev(dt_ctx *ctx, dt_mlp *m, const float *X, int64_t rows, float *Z) {
    DT_TRY(mlp_reserve(ctx, m, rows));
    return mlp_forward_internal(ctx, m, X, rows, Z);
}

extern "C" int dt_mlp_forward(dt_ctx *ctx, dt_mlp *m, const float *X, int64_t rows, float *Z) {
    DT_CHECK_ARG(ctx && m && X && Z && rows >= 1, "dt_mlp_forward: bad argument");
    return dt_mlp_forward_dev(ctx, m, X, rows, Z);
}

extern "C" int dt_mlp_loss_grad(dt_ctx *ctx, dt_mlp *m, const float *X, const float *T, int64_t rows,
                                double *loss, float *grad_host) {
    DT_CHECK_ARG(ctx && m && X && T && rows >= 1, "dt_mlp_loss_grad: bad argument");
    DT_TRY(mlp_reserve(ctx, m, rows));
    const int64_t H = m->hid, Do = m->dout;
    DT_TRY(mlp_forward_internal(ctx, m, X, rows, m->Z));
    const int64_t nz = rows * Do;
    const dim3 lg = dim3((unsigned)std::min<int64_t>((nz + 255) / 256, 2048));
    DT_LAUNCH("mlp_loss", mse_grad, lg, dim3(256), 0, m->Z, T, nz, 1.f / (float)rows, m->DZ, m->lpart);
    Epi none{nullptr, 0, nullptr, 0, nullptr, 0};
    // layer 3: dW3 = H2ᵀ dZ, db3 = colsum dZ, dH2 = (dZ W3ᵀ) * (H2 > 0)
    DT_TRY(mlp_gemm_launch(ctx, true, false, m->H2, H, m->DZ, Do, m->G + m->oW3, Do, H, Do, rows, none));
    DT_TRY(colsum(ctx, m, m->DZ, rows, Do, m->G + m->ob3));
    Epi mk2{nullptr, 0, m->H2, H, nullptr, 0};
    DT_TRY(mlp_gemm_launch(ctx, false, true, m->DZ, Do, m->P + m->oW3, Do, m->D2, H, rows, H, Do, mk2));
    // layer 2
    DT_TRY(mlp_gemm_launch(ctx, true, false, m->H1, H, m->D2, H, m->G + m->oW2, H, H, H, rows, none));
    DT_TRY(colsum(ctx, m, m->D2, rows, H, m->G + m->ob2));
    Epi mk1{nullptr, 0, m->H1, H, nullptr, 0};
    DT_TRY(mlp_gemm_launch(ctx, false, true, m->D2, H, m->P + m->oW2, H, m->D1, H, rows, H, H, mk1));
    // layer 1
    DT_TRY(mlp_gemm_launch(ctx, true, false, X, m->din, m->D1, H, m->G + m->oW1, H, m->din, H, rows, none));
    DT_TRY(colsum(ctx, m, m->D1, rows, H, m->G + m->ob1));
    if (loss) {
        std::vector<double> lp(lg.x);
        DT_HIP(hipMemcpyAsync(lp.data(), m->lpart, lg.x * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        DT_HIP(hipStreamSynchronize(ctx->stream));
        double s = 0.0;
        for (double v : lp) s += v;
        *loss = s / (double)rows;
    }
    if (grad_host) {
        DT_HIP(hipMemcpyAsync(grad_host, m->G, (size_t)m->nparam * 4, hipMemcpyDeviceToHost, ctx->stream));
        DT_HIP(hipStreamSynchronize(ctx->stream));
    }
    return DT_OK;
}

extern "C" int dt_mlp_adam_step(dt_ctx *ctx, dt_mlp *m, float lr) {
    DT_CHECK_ARG(ctx && m && lr > 0.f, "dt_mlp_adam_step: bad argument");
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    DT_LAUNCH("mlp_adam", adam_kernel, ew_grid(m->nparam), dim3(256), 0, m->P, m->G, m->Mo, m->Ve, m->nparam, lr,
              b1, b2, 1e-8f, c1, c2);
    return DT_OK;
}
