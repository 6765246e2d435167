// CPU restatement of the reference's dense HOOI-ALS in C++/OpenMP, FP64 --
// TEST INFRASTRUCTURE (the timed CPU baseline and a checker); the product
// never calls it.
//
// Follows Algorithm::Tucker::HOOI_ALS (template/algorithm/tucker/hooi_als.tpp:
// 64-123) on one rank: per sweep and mode n, Y = A.copy() (:96), then
// Y = ttm(Y, U_iᵀ, i) for i != n ascending (:97-101), U_n = ALS_(Y, n, U_n)
// (:103, five inner steps, :7-34), then the core G = A x_1 U_1ᵀ ... x_N U_Nᵀ
// (:105-109) and the residual sqrt(1 - ||G||^2/||A||^2) (:112-114).  Each
// ttm is the reference's unfold copy -> GEMM -> fold (Function::ttm,
// template/function/tensor.tpp:272-335: Operator::tenmatt, matmulNT,
// mattten, operator_cpu.tpp:83-135).  ALS_ uses ttt_except (tensor.tpp:
// 132-259), inverse (LU in the reference; Gauss-Jordan here) and reduced_QR
// (Householder in the reference; two passes of modified Gram-Schmidt here:
// the same span).
//
// The GEMMs are this file's own blocked kernels (no BLAS on the box): a packed
// 8 x 6 register-tiled DGEMM (AVX2 vector extensions) for general shapes and a
// streaming kernel for the tall-skinny unfold products; the unfold copies and
// A.copy() run over the OpenMP threads (with one MPI rank per core the
// reference's copies would run in parallel too).
//
// Modes:
//   cpu_hooi bench --shape 256,256,256,256 --ranks 16,16,16,16 [--sweeps 1] [--dgemm 4096]
//       synthetic Gaussian tensor; prints one JSON line: host, threads, DGEMM
//       GFLOP/s at n^3, start and per-sweep seconds, parts (copies, GEMMs, ALS).
//   cpu_hooi pieces --shape 256,256,256,256 --ranks 16,16,16,16
//       the two kinds of chain of a sweep timed at their real size as the
//       sweep runs them (A.copy() allocating afresh, as the reference's does)
//       and the sweep composed from them, for when a whole sweep would run too
//       long; prints one JSON line.
//   cpu_hooi check --shape I0,I1,.. --ranks R0,.. --sweeps S --tensor A.bin --start U.bin
//       A (column-major FP64) and the start factors (I_n x R_n column-major,
//       concatenated) from files; prints the residual of every sweep.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

typedef double v4d __attribute__((vector_size(32)));

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Parts {
    double copy = 0, unfold = 0, gemm = 0, fold = 0, als = 0;
};
Parts g_parts;

// ---- GEMM --------------------------------------------------------------------
// C (m x n, ldc) = A (m x k, lda) B (k x n, ldb) + beta C, column-major, with
// A or B read transposed when tA / tB.
constexpr int MR = 8, NR = 6, MC = 96, KC = 256, NC = 2016;

inline double at(const double *X, size_t ld, bool t, size_t i, size_t j) { return t ? X[j + ld * i] : X[i + ld * j]; }

void pack_a(const double *A, size_t lda, bool tA, size_t i0, size_t p0, size_t mc, size_t kc, double *Ap) {
    for (size_t ir = 0; ir < mc; ir += MR)
        for (size_t p = 0; p < kc; p++)
            for (int r = 0; r < MR; r++)
                *Ap++ = ir + r < mc ? at(A, lda, tA, i0 + ir + r, p0 + p) : 0.0;
}
void pack_b(const double *B, size_t ldb, bool tB, size_t p0, size_t j0, size_t kc, size_t nc, double *Bp) {
    for (size_t jr = 0; jr < nc; jr += NR)
        for (size_t p = 0; p < kc; p++)
            for (int c = 0; c < NR; c++)
                *Bp++ = jr + c < nc ? at(B, ldb, tB, p0 + p, j0 + jr + c) : 0.0;
}
// C[0:MR, 0:NR] (+)= Ap (kc x MR panel) Bp (kc x NR panel)
void micro(size_t kc, const double *Ap, const double *Bp, double *C, size_t ldc, size_t mr, size_t nr) {
    v4d c[NR][2];
    for (int j = 0; j < NR; j++) c[j][0] = c[j][1] = v4d{0, 0, 0, 0};
    for (size_t p = 0; p < kc; p++) {
        v4d a0, a1;
        std::memcpy(&a0, Ap + MR * p, 32);
        std::memcpy(&a1, Ap + MR * p + 4, 32);
        const double *b = Bp + NR * p;
        for (int j = 0; j < NR; j++) {
            const v4d bj = {b[j], b[j], b[j], b[j]};
            c[j][0] += a0 * bj;
            c[j][1] += a1 * bj;
        }
    }
    for (size_t j = 0; j < nr; j++)
        for (size_t i = 0; i < mr; i++) C[i + ldc * j] += c[j][i / 4][i % 4];
}

void dgemm(bool tA, bool tB, size_t m, size_t n, size_t k, const double *A, size_t lda, const double *B, size_t ldb,
           double beta, double *C, size_t ldc) {
    if (beta == 0.0) {
#pragma omp parallel for schedule(static)
        for (size_t j = 0; j < n; j++) std::memset(C + ldc * j, 0, m * sizeof(double));
    } else if (beta != 1.0) {
#pragma omp parallel for schedule(static)
        for (size_t j = 0; j < n; j++)
            for (size_t i = 0; i < m; i++) C[i + ldc * j] *= beta;
    }
    if (m == 0 || n == 0 || k == 0) return;
    std::vector<double> Bp((size_t)KC * (NC + NR));
    for (size_t j0 = 0; j0 < n; j0 += NC) {
        const size_t nc = std::min<size_t>(NC, n - j0);
        for (size_t p0 = 0; p0 < k; p0 += KC) {
            const size_t kc = std::min<size_t>(KC, k - p0);
            pack_b(B, ldb, tB, p0, j0, kc, nc, Bp.data());
#pragma omp parallel
            {
                std::vector<double> Ap((size_t)MC * KC + MR * KC);
#pragma omp for schedule(dynamic)
                for (size_t i0 = 0; i0 < m; i0 += MC) {
                    const size_t mc = std::min<size_t>(MC, m - i0);
                    pack_a(A, lda, tA, i0, p0, mc, kc, Ap.data());
                    for (size_t jr = 0; jr < nc; jr += NR)
                        for (size_t ir = 0; ir < mc; ir += MR)
                            micro(kc, Ap.data() + ir * kc, Bp.data() + jr * kc, C + (i0 + ir) + ldc * (j0 + jr), ldc,
                                  std::min<size_t>(MR, mc - ir), std::min<size_t>(NR, nc - jr));
                }
            }
        }
    }
}

// C (m x J) = X (m x k, column-major) M^T (M: J x k, J <= 64): the
// reference's matmulNT of an unfolding by a factor; X is streamed once
void gemm_tall_nt(size_t m, size_t J, size_t k, const double *X, const double *M, double *C) {
    constexpr size_t RB = 64;
#pragma omp parallel
    {
        std::vector<double> acc(RB * J);
#pragma omp for schedule(static)
        for (size_t r0 = 0; r0 < m; r0 += RB) {
            const size_t rb = std::min(RB, m - r0);
            std::fill(acc.begin(), acc.end(), 0.0);
            for (size_t i = 0; i < k; i++) {
                const double *x = X + r0 + m * i;
                for (size_t j = 0; j < J; j++) {
                    const double mj = M[j + J * i];
                    double *a = acc.data() + RB * j;
#pragma omp simd
                    for (size_t r = 0; r < rb; r++) a[r] += x[r] * mj;
                }
            }
            for (size_t j = 0; j < J; j++) std::memcpy(C + r0 + m * j, acc.data() + RB * j, rb * sizeof(double));
        }
    }
}

// ---- tensors (column-major, first index fastest) ------------------------------
struct Ten {
    std::vector<size_t> shape;
    std::vector<double> d;
    size_t size() const { return d.size(); }
};
void split(const std::vector<size_t> &sh, size_t n, size_t &F, size_t &I, size_t &B) {
    F = 1;
    B = 1;
    for (size_t i = 0; i < n; i++) F *= sh[i];
    for (size_t i = n + 1; i < sh.size(); i++) B *= sh[i];
    I = sh[n];
}
// dst[f + F q + F Q p] = src[f + F p + F P q] (f < F, p < P, q < Q): the
// unfold copies swap the middle and last index groups.  Runs of F >= 32
// doubles are copied whole; shorter runs go through 32 x 32 tiles so that
// both sides are read and written in runs.
void swap23(const double *src, size_t F, size_t P, size_t Q, double *dst) {
    if (F >= 32) {
#pragma omp parallel for collapse(2) schedule(static)
        for (size_t q = 0; q < Q; q++)
            for (size_t p = 0; p < P; p++) std::memcpy(dst + F * q + F * Q * p, src + F * p + F * P * q, F * 8);
        return;
    }
    constexpr size_t T = 32;
#pragma omp parallel
    {
        std::vector<double> tile(T * T * F);
#pragma omp for collapse(2) schedule(static)
        for (size_t q0 = 0; q0 < Q; q0 += T)
            for (size_t p0 = 0; p0 < P; p0 += T) {
                const size_t nq = std::min(T, Q - q0), np = std::min(T, P - p0);
                for (size_t q = 0; q < nq; q++)  // read runs of np F-chunks
                    std::memcpy(tile.data() + F * T * q, src + F * p0 + F * P * (q0 + q), F * np * 8);
                for (size_t p = 0; p < np; p++)  // write runs of nq F-chunks
                    for (size_t q = 0; q < nq; q++)
                        std::memcpy(dst + F * (q0 + q) + F * Q * (p0 + p), tile.data() + F * T * q + F * p, F * 8);
            }
    }
}
// tenmatt: Xt (rest x I_n), row r = f + F b (operator_cpu.tpp:83-103)
void tenmatt(const Ten &A, size_t n, std::vector<double> &Xt) {
    size_t F, I, B;
    split(A.shape, n, F, I, B);
    Xt.resize(A.size());
    swap23(A.d.data(), F, I, B, Xt.data());
}
// mattten: Y (shape with I_n -> J) from C (rest x J) (operator_cpu.tpp:115-135)
void mattten(const std::vector<double> &C, const std::vector<size_t> &sh, size_t n, size_t J, Ten &Y) {
    size_t F, I, B;
    split(sh, n, F, I, B);
    Y.shape = sh;
    Y.shape[n] = J;
    Y.d.resize(F * J * B);
    swap23(C.data(), F, B, J, Y.d.data());
}
// Function::ttm(A, M, n): Y = A x_n M, M J x I_n column-major
void ttm(const Ten &A, const std::vector<double> &M, size_t J, size_t n, Ten &Y, std::vector<double> &Xt,
         std::vector<double> &C) {
    size_t F, I, B;
    split(A.shape, n, F, I, B);
    const size_t rest = F * B;
    double t = now();
    tenmatt(A, n, Xt);
    g_parts.unfold += now() - t;
    t = now();
    C.resize(rest * J);
    if (J <= 64)
        gemm_tall_nt(rest, J, I, Xt.data(), M.data(), C.data());
    else
        dgemm(false, true, rest, J, I, Xt.data(), rest, M.data(), J, 0.0, C.data(), rest);
    g_parts.gemm += now() - t;
    t = now();
    mattten(C, A.shape, n, J, Y);
    g_parts.fold += now() - t;
}
void copy_ten(const Ten &A, Ten &Y) {
    const double t = now();
    Y.shape = A.shape;
    Y.d.resize(A.size());
    const size_t n = A.size();
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; c += (1 << 20)) std::memcpy(Y.d.data() + c, A.d.data() + c, std::min<size_t>(1 << 20, n - c) * 8);
    g_parts.copy += now() - t;
}
double sumsq(const std::vector<double> &v) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t i = 0; i < v.size(); i++) s += v[i] * v[i];
    return s;
}

// ---- small dense ----------------------------------------------------------------
// Q (m x n) <- orthonormal basis of span(Q): modified Gram-Schmidt, twice
void mgs2(std::vector<double> &Q, size_t m, size_t n) {
    for (int pass = 0; pass < 2; pass++)
        for (size_t j = 0; j < n; j++) {
            double *qj = Q.data() + m * j;
            for (size_t i = 0; i < j; i++) {
                const double *qi = Q.data() + m * i;
                double d = 0.0;
                for (size_t r = 0; r < m; r++) d += qi[r] * qj[r];
                for (size_t r = 0; r < m; r++) qj[r] -= d * qi[r];
            }
            double nr = 0.0;
            for (size_t r = 0; r < m; r++) nr += qj[r] * qj[r];
            nr = std::sqrt(nr);
            for (size_t r = 0; r < m; r++) qj[r] /= nr;
        }
}
// in-place inverse (Gauss-Jordan, partial pivoting) of n x n column-major
void inverse(std::vector<double> &A, size_t n) {
    std::vector<size_t> perm(n);
    for (size_t k = 0; k < n; k++) {
        size_t p = k;
        for (size_t i = k + 1; i < n; i++)
            if (std::fabs(A[i + n * k]) > std::fabs(A[p + n * k])) p = i;
        perm[k] = p;
        if (p != k)
            for (size_t j = 0; j < n; j++) std::swap(A[k + n * j], A[p + n * j]);
        const double piv = 1.0 / A[k + n * k];
        A[k + n * k] = 1.0;
        for (size_t j = 0; j < n; j++) A[k + n * j] *= piv;
        for (size_t i = 0; i < n; i++) {
            if (i == k) continue;
            const double f = A[i + n * k];
            A[i + n * k] = 0.0;
            for (size_t j = 0; j < n; j++) A[i + n * j] -= f * A[k + n * j];
        }
    }
    for (size_t k = n; k-- > 0;)
        if (perm[k] != k)
            for (size_t i = 0; i < n; i++) std::swap(A[i + n * k], A[i + n * perm[k]]);
}

// ALS_ (hooi_als.tpp:7-34) on Y (mode n of size I), L_initial I x R
void als(const Ten &Y, size_t n, std::vector<double> &L, size_t R, int inner) {
    const double t0 = now();
    size_t F, I, B;
    split(Y.shape, n, F, I, B);
    std::vector<double> P(I * R), Gm(R * R), W(I * R), GR(R * R), Pt(R * I), Xt, C;
    Ten Z;
    // Y_(n) as I x rest (tenmat) for ttt_except
    std::vector<double> Yt;
    tenmatt(Y, n, Yt);  // rest x I = Y_(n)^T
    const size_t rest = F * B;
    for (int it = 0; it < inner; it++) {
        if (it == 0) {
            P = L;
        } else {
            dgemm(true, false, R, R, I, L.data(), I, L.data(), I, 0.0, Gm.data(), R);  // L^T L
            inverse(Gm, R);
            dgemm(false, false, I, R, R, L.data(), I, Gm.data(), R, 0.0, P.data(), I);  // P = L (L^T L)^{-1}
        }
        // W = ttt_except(Y, Y x_n P^T, n) = Y_(n) (Y x_n P^T)_(n)^T
        for (size_t i = 0; i < I; i++)
            for (size_t r = 0; r < R; r++) Pt[r + R * i] = P[i + I * r];
        ttm(Y, Pt, R, n, Z, Xt, C);
        std::vector<double> Zx;
        tenmatt(Z, n, Zx);  // rest x R
        dgemm(true, false, I, R, rest, Yt.data(), rest, Zx.data(), rest, 0.0, W.data(), I);  // Y_(n) Z_(n)^T
        dgemm(true, false, R, R, I, P.data(), I, W.data(), I, 0.0, GR.data(), R);           // G_R = P^T W
        inverse(GR, R);
        dgemm(false, false, I, R, R, W.data(), I, GR.data(), R, 0.0, L.data(), I);  // L = W G_R^{-1}
    }
    mgs2(L, I, R);  // reduced_QR's Q (hooi_als.tpp:32)
    g_parts.als += now() - t0;
}

struct Hooi {
    std::vector<size_t> I, R;
    std::vector<std::vector<double>> U;  // I_n x R_n
};

// one sweep (hooi_als.tpp:92-115); returns the residual
double sweep(const Ten &A, double normA2, Hooi &h, int inner) {
    const size_t N = A.shape.size();
    Ten Y, T;
    std::vector<double> Xt, C, Mt;
    auto factor_t = [&](size_t i) {  // U_i^T (R_i x I_i), the reference's transpose (:99)
        Mt.resize(h.R[i] * h.I[i]);
        for (size_t a = 0; a < h.I[i]; a++)
            for (size_t r = 0; r < h.R[i]; r++) Mt[r + h.R[i] * a] = h.U[i][a + h.I[i] * r];
    };
    for (size_t n = 0; n < N; n++) {
        copy_ten(A, Y);
        for (size_t i = 0; i < N; i++) {
            if (i == n) continue;
            factor_t(i);
            ttm(Y, Mt, h.R[i], i, T, Xt, C);
            std::swap(Y, T);
        }
        als(Y, n, h.U[n], h.R[n], inner);
    }
    copy_ten(A, Y);
    for (size_t i = 0; i < N; i++) {
        factor_t(i);
        ttm(Y, Mt, h.R[i], i, T, Xt, C);
        std::swap(Y, T);
    }
    const double g2 = sumsq(Y.d);
    return std::sqrt(std::max(0.0, 1.0 - g2 / normA2));
}

std::vector<size_t> parse_list(const char *s) {
    std::vector<size_t> v;
    const char *p = s;
    while (*p) {
        char *e = nullptr;
        v.push_back(std::strtoull(p, &e, 10));
        p = *e == ',' ? e + 1 : e;
        if (!*e) break;
    }
    return v;
}

uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
// Gaussian fill, a counter-based stream per 2^20 block (thread-count independent)
void fill_normal(std::vector<double> &v, uint64_t seed) {
    const size_t n = v.size(), bs = (size_t)1 << 20;
#pragma omp parallel for schedule(static)
    for (size_t b0 = 0; b0 < n; b0 += bs) {
        uint64_t s = seed ^ (b0 * 0x2545F4914F6CDD1DULL);
        for (size_t i = b0; i < std::min(n, b0 + bs); i += 2) {
            const double u1 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            const double u2 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            const double r = std::sqrt(-2.0 * std::log(u1));
            v[i] = r * std::cos(6.283185307179586 * u2);
            if (i + 1 < n) v[i + 1] = r * std::sin(6.283185307179586 * u2);
        }
    }
}

std::string cpu_model() {
    std::ifstream f("/proc/cpuinfo");
    std::string line;
    while (std::getline(f, line))
        if (line.rfind("model name", 0) == 0) {
            const size_t c = line.find(':');
            return c == std::string::npos ? line : line.substr(c + 2);
        }
    return "unknown";
}

const char *arg(int argc, char **argv, const char *key, const char *def) {
    for (int i = 2; i + 1 < argc; i++)
        if (std::strcmp(argv[i], key) == 0) return argv[i + 1];
    return def;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s bench|check --shape ... --ranks ... [options]\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    Hooi h;
    h.I = parse_list(arg(argc, argv, "--shape", "64,64,64"));
    h.R = parse_list(arg(argc, argv, "--ranks", "8,8,8"));
    const int sweeps = std::atoi(arg(argc, argv, "--sweeps", "1"));
    const int inner = std::atoi(arg(argc, argv, "--inner", "5"));
    if (h.I.size() != h.R.size() || h.I.empty()) {
        std::fprintf(stderr, "shape and ranks differ in length\n");
        return 2;
    }
    const size_t N = h.I.size();
    Ten A;
    A.shape = h.I;
    size_t total = 1;
    for (size_t x : h.I) total *= x;
    A.d.resize(total);
    h.U.resize(N);
    if (mode == "check") {
        std::ifstream ft(arg(argc, argv, "--tensor", ""), std::ios::binary), fu(arg(argc, argv, "--start", ""), std::ios::binary);
        if (!ft || !fu) {
            std::fprintf(stderr, "cannot open --tensor / --start\n");
            return 2;
        }
        ft.read((char *)A.d.data(), total * 8);
        for (size_t n = 0; n < N; n++) {
            h.U[n].resize(h.I[n] * h.R[n]);
            fu.read((char *)h.U[n].data(), h.U[n].size() * 8);
        }
        const double normA2 = sumsq(A.d);
        for (int s = 0; s < sweeps; s++) std::printf("%.17g\n", sweep(A, normA2, h, inner));
        return 0;
    }
    // bench / pieces
    const int threads = omp_get_max_threads();
    const size_t dn = std::strtoull(arg(argc, argv, "--dgemm", "4096"), nullptr, 10);
    double gflops = 0.0;
    if (dn > 0) {
        std::vector<double> X(dn * dn), Yd(dn * dn), Z(dn * dn);
        fill_normal(X, 1);
        fill_normal(Yd, 2);
        dgemm(false, false, 256, 256, 256, X.data(), dn, Yd.data(), dn, 0.0, Z.data(), dn);  // warm-up
        const double t = now();
        dgemm(false, false, dn, dn, dn, X.data(), dn, Yd.data(), dn, 0.0, Z.data(), dn);
        gflops = 2.0 * dn * dn * dn / (now() - t) / 1e9;
    }
    double t = now();
    fill_normal(A.d, 20000905);
    for (size_t n = 0; n < N; n++) {
        h.U[n].resize(h.I[n] * h.R[n]);
        fill_normal(h.U[n], 100 + n);
        mgs2(h.U[n], h.I[n], h.R[n]);
    }
    const double normA2 = sumsq(A.d);
    const double gen_s = now() - t;
    std::string shape;
    for (size_t n = 0; n < N; n++) shape += (n ? "," : "") + std::to_string(h.I[n]);
    if (mode == "pieces") {
        // One sweep of the reference schedule is N + 1 chains, each A.copy()
        // and products over the other modes; chains 1 .. N-1 and the core chain
        // start with a product over mode 0, chain 0 over mode 1.  Chains 0 and
        // 1 (with their ALS_ updates) are timed as the sweep runs them, and the
        // sweep is composed as chain 0 + N x chain 1 (the core chain's extra
        // product on an R^N-sized tensor counted as chain 1's ALS_ update).
        double chain_s[2];
        Ten Y, T;
        std::vector<double> Xt, C, Mt;  // kept across chains, as in sweep()
        for (size_t n = 0; n < 2 && n < N; n++) {
            g_parts = Parts();
            const double t0 = now();
            copy_ten(A, Y);
            for (size_t i = 0; i < N; i++) {
                if (i == n) continue;
                Mt.resize(h.R[i] * h.I[i]);
                for (size_t a = 0; a < h.I[i]; a++)
                    for (size_t r = 0; r < h.R[i]; r++) Mt[r + h.R[i] * a] = h.U[i][a + h.I[i] * r];
                ttm(Y, Mt, h.R[i], i, T, Xt, C);
                std::swap(Y, T);
            }
            als(Y, n, h.U[n], h.R[n], inner);
            chain_s[n] = now() - t0;
        }
        const double sweep_est = chain_s[0] + (double)N * chain_s[N > 1 ? 1 : 0];
        std::printf("{\"host\": \"%s\", \"nproc\": %d, \"threads\": %d, \"dgemm_n\": %zu, \"dgemm_gflops\": %.2f, "
                    "\"shape\": [%s], \"gen_s\": %.3f, \"chain_s\": [%.4f, %.4f], \"sweep_s_composed\": %.4f, "
                    "\"parts_chain1_s\": {\"copy\": %.4f, \"unfold\": %.4f, \"gemm\": %.4f, \"fold\": %.4f, \"als\": %.4f}}\n",
                    cpu_model().c_str(), omp_get_num_procs(), threads, dn, gflops, shape.c_str(), gen_s, chain_s[0],
                    chain_s[N > 1 ? 1 : 0], sweep_est, g_parts.copy, g_parts.unfold, g_parts.gemm, g_parts.fold,
                    g_parts.als);
        return 0;
    }
    std::vector<double> sw;
    double res = 0.0;
    for (int s = 0; s < sweeps; s++) {
        g_parts = Parts();
        t = now();
        res = sweep(A, normA2, h, inner);
        sw.push_back(now() - t);
    }
    std::printf("{\"host\": \"%s\", \"nproc\": %d, \"threads\": %d, \"dgemm_n\": %zu, \"dgemm_gflops\": %.2f, "
                "\"shape\": [%s], \"gen_s\": %.3f, \"sweep_s\": [",
                cpu_model().c_str(), omp_get_num_procs(), threads, dn, gflops, shape.c_str(), gen_s);
    for (size_t i = 0; i < sw.size(); i++) std::printf("%s%.4f", i ? ", " : "", sw[i]);
    std::printf("], \"residual\": %.9f, \"parts_last_sweep_s\": {\"copy\": %.4f, \"unfold\": %.4f, \"gemm\": %.4f, "
                "\"fold\": %.4f, \"als\": %.4f}}\n",
                res, g_parts.copy, g_parts.unfold, g_parts.gemm, g_parts.fold, g_parts.als);
    return 0;
}
