// Learned start factors: Gaussian sketch -> features -> row-wise MLP -> QR.
//
// No counterpart in the reference, which starts HOOI from random factors
// (template/algorithm/tucker/hooi_als.tpp:76-81).  For mode n the sketch is
// Y = A_(n) Ω with Ω a Gaussian (prod_{m!=n} I_m) x S matrix stored as a
// tensor of A's shape with mode n replaced by S (Philox stream 5000 + n), so
// Y is the cross product A_(n) Ω_(n)ᵀ (dt_cross_internal, split-K MFMA).
// The Kronecker sketch instead takes Ω = ⊗_{m != n} Ω_m with Ω_m Gaussian
// I_m x q: Y_(n) = (A x_{m != n} Ω_mᵀ)_(n) is a chain of mode products on
// the streaming TTM kernels (one pass over A per mode), S = q^(N-1).
// With YᵀY = V diag(σ²) Vᵀ (σ descending), the features are
//     X = sqrt(I_n) / σ_1 · Y V          (I_n x S, row-major FP32),
// i.e. the sketch's left singular vectors scaled by σ_k / σ_1, each column's
// sign fixed so that its entry of largest magnitude is positive.  The first
// R columns span the randomized range finder's estimate of the leading
// subspace; the MLP's residual form (z = x[:R] + ...) starts from exactly
// that estimate.
#include "dt_internal.hpp"

#include <cmath>
#include <vector>

int dt_mlp_forward_dev(dt_ctx *ctx, dt_mlp *m, const float *X, int64_t rows, float *Z);

namespace {

// sg[k] = sign of the entry of largest magnitude of column k of YV (first
// index on ties), so the features do not depend on the eigen-solver's signs.
__global__ __launch_bounds__(256) void column_signs(const double *__restrict__ YV, int64_t I,
                                                   double *__restrict__ sg) {
    __shared__ double bv[256];
    __shared__ int64_t bi[256];
    const int k = blockIdx.x;
    double best = -1.0;
    int64_t arg = 0;
    for (int64_t i = threadIdx.x; i < I; i += 256) {
        const double a = fabs(YV[i + I * k]);
        if (a > best) {
            best = a;
            arg = i;
        }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = arg;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const double v2 = bv[threadIdx.x + o];
            const int64_t i2 = bi[threadIdx.x + o];
            if (v2 > bv[threadIdx.x] || (v2 == bv[threadIdx.x] && i2 < bi[threadIdx.x])) {
                bv[threadIdx.x] = v2;
                bi[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) sg[k] = YV[bi[0] + I * k] < 0.0 ? -1.0 : 1.0;
}

// X[i * S + k] = sg[k] * scale * YV[i + I k], scale = sqrt(I) / sqrt(w[0])
__global__ void sketch_rows(const double *__restrict__ YV, const double *__restrict__ w,
                           const double *__restrict__ sg, int64_t I, int S, float *__restrict__ X) {
    const double s1 = sqrt(w[0]);
    const double scale = s1 > 0.0 ? sqrt((double)I) / s1 : 0.0;
    const int64_t total = I * S;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / S;
        const int k = (int)(e % S);
        X[e] = (float)(sg[k] * scale * YV[i + I * k]);
    }
}

// U (I x R, column-major FP64) = Z rows (row-major FP32, stride R)
__global__ void rows_to_cols(const float *__restrict__ Z, int64_t I, int R, double *__restrict__ U) {
    const int64_t total = I * R;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % I;
        const int j = (int)(e / I);
        U[e] = (double)Z[i * R + j];
    }
}

inline dim3 lg_grid(int64_t n) { return dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)); }

}  // namespace

// Feature pipeline of a sketch Y (I x S, column-major FP64): G = YᵀY, the
// eigenvectors V of G (Jacobi), X = signs * sqrt(I) / sqrt(w_0) * Y V.  The
// Jacobi step of several modes runs as one launch (one workgroup per mode).
struct SketchWs {
    double *Y, *YV, *sg;  // I x S, I x S, S
};
static int sketch_gram(dt_ctx *ctx, const SketchWs &s, int64_t I, int S, double *G) {
    return dt_gemm_f64_internal(ctx, 1, 0, (size_t)S, (size_t)S, (size_t)I, 1.0, s.Y, (size_t)I, s.Y, (size_t)I, 0.0,
                                G, (size_t)S);
}
static int sketch_finish(dt_ctx *ctx, const SketchWs &s, int64_t I, int S, const double *V, const double *w,
                         float *X) {
    DT_TRY(dt_gemm_f64_internal(ctx, 0, 0, (size_t)I, (size_t)S, (size_t)S, 1.0, s.Y, (size_t)I, V, (size_t)S, 0.0,
                                s.YV, (size_t)I));
    DT_LAUNCH("sketch_rows", column_signs, dim3((unsigned)S), dim3(256), 0, s.YV, I, s.sg);
    DT_LAUNCH("sketch_rows", sketch_rows, lg_grid(I * S), dim3(256), 0, s.YV, w, s.sg, I, S, X);
    return DT_OK;
}

// X from the sketch Y (I x S, column-major FP64) held at the start of ws
// (which has room for 2 I S + 2 S^2 + 2 S doubles).
static int features_from_sketch(dt_ctx *ctx, double *ws, int64_t I, int S, float *X) {
    SketchWs s{ws, ws + I * S, nullptr};
    double *G = s.YV + I * S;
    double *V = G + (int64_t)S * S;
    double *w = V + (int64_t)S * S;
    s.sg = w + S;
    DT_TRY(sketch_gram(ctx, s, I, S, G));
    DT_TRY(dt_jacobi_eig(ctx, G, (size_t)S, V, w));
    return sketch_finish(ctx, s, I, S, V, w, X);
}

// Gaussian sketch of mode n (Ydst: as for sketch_kron)
static int sketch_gauss(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape, int n, int S,
                        uint64_t seed, float *X, double *Ydst) {
    DT_CHECK_ARG(ctx && A && shape && (X || Ydst), "dt_sketch_features: null argument");
    DT_CHECK_ARG(ndim >= 2 && ndim <= DT_MAX_NDIM && n >= 0 && n < ndim, "dt_sketch_features: bad mode");
    DT_CHECK_ARG(dtA == DT_F32 || dtA == DT_BF16, "dt_sketch_features: tensor must be F32 or BF16");
    DT_CHECK_ARG(S >= 1 && S <= 96, "dt_sketch_features: sketch width %d not in [1, 96]", S);
    int64_t F = 1, Bk = 1;
    for (int i = 0; i < n; i++) F *= (int64_t)shape[i];
    for (int i = n + 1; i < ndim; i++) Bk *= (int64_t)shape[i];
    const int64_t I = (int64_t)shape[n];
    const int64_t nom = F * S * Bk;
    const size_t esz = dtA == DT_BF16 ? 2 : 4;
    void *Om = dt_ws(ctx, WS_MLP0, (size_t)nom * esz);
    double *ws = (double *)dt_ws(ctx, WS_MLP1, (size_t)(2 * I * S + 2 * S * S + 2 * S + 64) * sizeof(double));
    if (!Om || !ws) return DT_ENOMEM;
    DT_TRY(dt_gen_normal(ctx, Om, dtA, (size_t)nom, seed, 5000 + (uint64_t)n));
    DT_TRY(dt_cross_internal(ctx, A, Om, dtA, F, I, S, Bk, Ydst ? Ydst : ws));
    if (Ydst) return DT_OK;
    return features_from_sketch(ctx, ws, I, S, X);
}

extern "C" int dt_sketch_features(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape, int n,
                                  int S, uint64_t seed, float *X) {
    return sketch_gauss(ctx, A, dtA, ndim, shape, n, S, seed, X, nullptr);
}

// Y_(n) (I x S column-major FP64) of a tensor T whose mode n has size I and
// every other mode size q: column = Kolda index of the remaining modes.
__global__ void unfold_to_f64(const float *__restrict__ T, int ndim, int n, int64_t I, int q, int64_t total,
                              double *__restrict__ Y) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t rem = e, col = 0, stride = 1, i = 0;
        for (int m = 0; m < ndim; m++) {
            if (m == n) {
                i = rem % I;
                rem /= I;
            } else {
                col += (rem % q) * stride;
                rem /= q;
                stride *= q;
            }
        }
        Y[i + I * col] = (double)T[e];
    }
}

// Kronecker sketch of mode n: with Ydst the unfolded sketch (I x S FP64) is
// written there and nothing else happens; without, the features go to X.
// pre (optional, modes 0 and 1 of a tensor of 3 or more modes): A already
// contracted over modes ndim - 1 .. 2 with their Ω in that order (FP32,
// I_0 x I_1 x q ...), the prefix both modes' chains share; the chain then
// runs only its last product, with the same kernels and operands, so the
// sketch is the same.
static int sketch_kron(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape, int n, int q,
                       uint64_t seed, float *X, double *Ydst, const float *pre = nullptr) {
    DT_CHECK_ARG(ctx && A && shape && (X || Ydst), "dt_sketch_features_kron: null argument");
    DT_CHECK_ARG(ndim >= 2 && ndim <= DT_MAX_NDIM && n >= 0 && n < ndim, "dt_sketch_features_kron: bad mode");
    DT_CHECK_ARG(dtA == DT_F32 || dtA == DT_BF16, "dt_sketch_features_kron: tensor must be F32 or BF16");
    int64_t S = 1;
    for (int m = 0; m < ndim - 1; m++) S *= q;
    DT_CHECK_ARG(q >= 1 && S <= 96, "dt_sketch_features_kron: q^(ndim-1) = %lld not in [1, 96]", (long long)S);
    const int64_t I = (int64_t)shape[n];
    // factors Ω_m (q x I_m, column-major FP32), Philox stream 6000 + m
    size_t off[DT_MAX_NDIM], tot = 0;
    for (int m = 0; m < ndim; m++) {
        off[m] = tot;
        tot += m == n ? 0 : (size_t)q * shape[m];
    }
    float *Om = (float *)dt_ws(ctx, WS_MLP0, tot * sizeof(float) + 256);
    double *ws = (double *)dt_ws(ctx, WS_MLP1, (size_t)(2 * I * S + 2 * S * S + 2 * S + 64) * sizeof(double) +
                                                   (size_t)(I * S) * sizeof(float));
    if (!Om || !ws) return DT_ENOMEM;
    float *Yt = (float *)(ws + 2 * I * S + 2 * S * S + 2 * S + 64);
    for (int m = 0; m < ndim; m++)
        if (m != n) DT_TRY(dt_gen_normal(ctx, Om + off[m], DT_F32, (size_t)q * shape[m], seed, 6000 + (uint64_t)m));
    // contraction order: the fused leading pair when n >= 2, else highest mode first
    int idx[DT_MAX_NDIM], nm = 0;
    if (n >= 2) {
        idx[nm++] = 0;
        idx[nm++] = 1;
        for (int m = ndim - 1; m >= 2; m--)
            if (m != n) idx[nm++] = m;
    } else {
        for (int m = ndim - 1; m >= 0; m--)
            if (m != n) idx[nm++] = m;
    }
    const void *Ms[DT_MAX_NDIM];
    size_t Js[DT_MAX_NDIM];
    for (int k = 0; k < nm; k++) {
        Ms[k] = Om + off[idx[k]];
        Js[k] = (size_t)q;
    }
    if (pre && n < 2 && ndim >= 3) {
        size_t psh[DT_MAX_NDIM];
        for (int m = 0; m < ndim; m++) psh[m] = m < 2 ? shape[m] : (size_t)q;
        DT_TRY(dt_ttmc(ctx, pre, DT_F32, ndim, psh, Ms + (nm - 1), DT_F32, Js + (nm - 1), idx + (nm - 1), 1, Yt,
                       DT_F32));
    } else {
        DT_TRY(dt_ttmc(ctx, A, dtA, ndim, shape, Ms, DT_F32, Js, idx, nm, Yt, DT_F32));
    }
    DT_LAUNCH("sketch_unfold", unfold_to_f64, lg_grid(I * S), dim3(256), 0, Yt, ndim, n, I, q, I * S,
              Ydst ? Ydst : ws);
    if (Ydst) return DT_OK;
    return features_from_sketch(ctx, ws, I, (int)S, X);
}

extern "C" int dt_sketch_features_kron(dt_ctx *ctx, const void *A, int dtA, int ndim, const size_t *shape, int n,
                                       int q, uint64_t seed, float *X) {
    return sketch_kron(ctx, A, dtA, ndim, shape, n, q, seed, X, nullptr);
}

extern "C" int dt_learned_start(dt_ctx *ctx, dt_mlp *mlp, const void *A, int dtA, int ndim, const size_t *shape,
                                const size_t *ranks, int sketch, uint64_t seed, double *const *U) {
    DT_CHECK_ARG(ctx && mlp && A && shape && ranks && U, "dt_learned_start: null argument");
    DT_CHECK_ARG(ndim >= 2 && ndim <= DT_MAX_NDIM, "dt_learned_start: bad ndim");
    int din = 0, hid = 0, dout = 0;
    int64_t np = 0;
    dt_mlp_shape(mlp, &din, &hid, &dout, &np);
    DT_CHECK_ARG(sketch == DT_SKETCH_GAUSSIAN || sketch == DT_SKETCH_KRONECKER, "dt_learned_start: bad sketch %d",
                 sketch);
    int q = 0;
    if (sketch == DT_SKETCH_KRONECKER) {
        q = (int)std::lround(std::pow((double)din, 1.0 / (ndim - 1)));
        int64_t S = 1;
        for (int m = 0; m < ndim - 1; m++) S *= q;
        DT_CHECK_ARG(S == din, "dt_learned_start: model input %d is not q^%d", din, ndim - 1);
    }
    int64_t rows = 0;
    for (int n = 0; n < ndim; n++) {
        DT_CHECK_ARG(U[n], "dt_learned_start: factor %d missing", n);
        DT_CHECK_ARG((int)ranks[n] == dout, "dt_learned_start: rank %zu of mode %d differs from the model's %d",
                     ranks[n], n, dout);
        DT_CHECK_ARG(ranks[n] <= shape[n], "dt_learned_start: rank above mode size");
        rows += (int64_t)shape[n];
    }
    float *X = (float *)dt_ws(ctx, WS_MLP2, (size_t)rows * (din + dout) * sizeof(float));
    // per mode: sketch Y and Y V (I_n x S each), signs (S); then all Grams, V, w
    const int S = din;
    const size_t per_mode_rows = (size_t)rows;
    double *sk = (double *)dt_ws(ctx, WS_SKETCH, (2 * per_mode_rows * S + (size_t)ndim * (2 * S * S + 2 * S) + 64) *
                                                    sizeof(double));
    if (!X || !sk) return DT_ENOMEM;
    float *Z = X + rows * din;
    SketchWs sw[DT_MAX_NDIM];
    double *G = sk + 2 * per_mode_rows * S;
    double *V = G + (size_t)ndim * S * S;
    double *w = V + (size_t)ndim * S * S;
    double *sg = w + (size_t)ndim * S;
    // Kronecker sketches of modes 0 and 1 share A x_{ndim-1} Ω ... x_2 Ω (a full
    // read of A): computed once here
    float *pre = nullptr;
    if (sketch == DT_SKETCH_KRONECKER && ndim >= 3) {
        size_t pe = shape[0] * shape[1], off_m[DT_MAX_NDIM], tot = 0;
        for (int m = 2; m < ndim; m++) {
            pe *= (size_t)q;
            off_m[m] = tot;
            tot += (size_t)q * shape[m];
        }
        float *Om = (float *)dt_ws(ctx, WS_MLP0, tot * sizeof(float) + 256);
        pre = (float *)dt_ws(ctx, WS_SKPRE, pe * sizeof(float));
        if (!Om || !pre) return DT_ENOMEM;
        const void *Ms[DT_MAX_NDIM];
        size_t Js[DT_MAX_NDIM];
        int idx[DT_MAX_NDIM], nm = 0;
        for (int m = ndim - 1; m >= 2; m--) {
            DT_TRY(dt_gen_normal(ctx, Om + off_m[m], DT_F32, (size_t)q * shape[m], seed, 6000 + (uint64_t)m));
            Ms[nm] = Om + off_m[m];
            Js[nm] = (size_t)q;
            idx[nm++] = m;
        }
        DT_TRY(dt_ttmc(ctx, A, dtA, ndim, shape, Ms, DT_F32, Js, idx, nm, pre, DT_F32));
    }
    int64_t off = 0;
    for (int n = 0; n < ndim; n++) {
        const int64_t I = (int64_t)shape[n];
        sw[n] = SketchWs{sk + 2 * off * S, sk + 2 * off * S + I * S, sg + (size_t)n * S};
        if (sketch == DT_SKETCH_KRONECKER)
            DT_TRY(sketch_kron(ctx, A, dtA, ndim, shape, n, q, seed, nullptr, sw[n].Y, pre));
        else
            DT_TRY(sketch_gauss(ctx, A, dtA, ndim, shape, n, S, seed, nullptr, sw[n].Y));
        DT_TRY(sketch_gram(ctx, sw[n], I, S, G + (size_t)n * S * S));
        off += I;
    }
    // the eigen-decompositions of all modes in one launch (a few workgroups:
    // queued side work, such as the tensor norms, runs beside it)
    DT_TRY(dt_side_kick(ctx));
    DT_TRY(dt_jacobi_eig_batched(ctx, G, (size_t)S, ndim, V, w));
    off = 0;
    for (int n = 0; n < ndim; n++) {
        const int64_t I = (int64_t)shape[n];
        DT_TRY(sketch_finish(ctx, sw[n], I, S, V + (size_t)n * S * S, w + (size_t)n * S, X + off * din));
        off += I;
    }
    DT_TRY(dt_mlp_forward_dev(ctx, mlp, X, rows, Z));
    off = 0;
    for (int n = 0; n < ndim; n++) {
        const int64_t I = (int64_t)shape[n];
        DT_LAUNCH("learned_rows", rows_to_cols, lg_grid(I * dout), dim3(256), 0, Z + off * dout, I, dout, U[n]);
        DT_TRY(dt_qr_f64(ctx, U[n], nullptr, U[n], (size_t)I, (size_t)dout));
        off += I;
    }
    return DT_OK;
}

// HOOI hook (dt_hooi.hip): learned start for every mode.
int dt_learned_init(dt_ctx *ctx, const void *model, const void *A, int dtA, int N, const int64_t *I,
                    const int64_t *R, int sketch, uint64_t seed, double *const *U64) {
    if (!model) return dt_fail(ctx, DT_EINVAL, "learned init: no model (dt_hooi_args.model)");
    size_t sh[DT_MAX_NDIM], rk[DT_MAX_NDIM];
    for (int n = 0; n < N; n++) {
        sh[n] = (size_t)I[n];
        rk[n] = (size_t)R[n];
    }
    return dt_learned_start(ctx, (dt_mlp *)model, A, dtA, N, sh, rk, sketch, seed, U64);
}
