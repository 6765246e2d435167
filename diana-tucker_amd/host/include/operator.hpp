// Operator<Ty>: the reference's raw-pointer, column-major, destination-first
// compute primitives (include/operator.hpp:8-70).  Same names, argument order
// and meaning; host pointers in and out, the arithmetic runs in
// libdiana_hip.so on the GPU (FP64 for the dense small-matrix kernels, as the
// reference's LAPACK/BLAS calls in src/operator/operator_double.cpp).
// Elementwise helpers and layout maps stay on the host, as in the reference.
#ifndef DIANA_AMD_OPERATOR_HPP
#define DIANA_AMD_OPERATOR_HPP
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "def.hpp"
#include "device.hpp"
#include "util.hpp"

template <typename Ty>
class Operator {
    static std::vector<double> to_f64(const Ty *A, size_t n) { return std::vector<double>(A, A + n); }
    static void from_f64(Ty *dst, const std::vector<double> &v) {
        for (size_t i = 0; i < v.size(); i++) dst[i] = (Ty)v[i];
    }

public:
    static Ty *alloc(size_t n) { return (Ty *)std::malloc(n * sizeof(Ty)); }
    static void free(Ty *A) { std::free(A); }
    static void mcpy(Ty *dest, Ty *src, size_t len) { std::memcpy(dest, src, len * sizeof(Ty)); }
    static void add(Ty *C, Ty *A, Ty *B, size_t n) { for (size_t i = 0; i < n; i++) C[i] = A[i] + B[i]; }
    static void sub(Ty *C, Ty *A, Ty *B, size_t n) { for (size_t i = 0; i < n; i++) C[i] = A[i] - B[i]; }
    static void mul(Ty *C, Ty *A, Ty *B, size_t n) { for (size_t i = 0; i < n; i++) C[i] = A[i] * B[i]; }
    static void nmul(Ty *C, Ty *A, Ty b, size_t n) { for (size_t i = 0; i < n; i++) C[i] = A[i] * b; }
    static void eye(Ty *A, size_t M) {
        for (size_t i = 0; i < M * M; i++) A[i] = 0;
        for (size_t i = 0; i < M; i++) A[i * M + i] = 1;
    }
    static void constant(Ty *A, Ty c, size_t n) { for (size_t i = 0; i < n; i++) A[i] = c; }
    // operator_double.cpp:58-72, single-threaded so the sequence is deterministic
    static void rand(Ty *A, size_t n) { for (size_t i = 0; i < n; i++) A[i] = (Ty)((double)std::rand() / RAND_MAX); }
    static void randn(Ty *A, size_t n) { for (size_t i = 0; i < n; i++) A[i] = (Ty)Util::randn(); }

    static void inverse(Ty *C, Ty *A, size_t n) {
        auto a = to_f64(A, n * n);
        diana::DevBuf dA(a.data(), a.size() * 8), dC(a.size() * 8);
        diana::check(dt_inverse_f64(diana::Device::ctx(), (double *)dC.p, (double *)dA.p, n), "inverse");
        dC.to_host(a.data(), a.size() * 8);
        from_f64(C, a);
    }
    // Reduced QR: Q m x n, R n x n (operator_double.cpp:131-155).
    static void QR(Ty *Q, Ty *R, Ty *A, size_t m, size_t n) {
        auto a = to_f64(A, m * n);
        std::vector<double> r(n * n);
        diana::DevBuf dA(a.data(), a.size() * 8), dQ(a.size() * 8), dR(r.size() * 8);
        diana::check(dt_qr_f64(diana::Device::ctx(), (double *)dQ.p, (double *)dR.p, (double *)dA.p, m, n), "QR");
        dQ.to_host(a.data(), a.size() * 8);
        dR.to_host(r.data(), r.size() * 8);
        from_f64(Q, a);
        from_f64(R, r);
    }
    // LQ of an m x n (m <= n) matrix (operator_double.cpp:105-129): A = L Q.
    static void LQ(Ty *L, Ty *Q, Ty *A, size_t m, size_t n) {
        std::vector<double> at(m * n), q(m * n), r(m * m);
        for (size_t i = 0; i < m; i++)
            for (size_t j = 0; j < n; j++) at[j + n * i] = (double)A[i + m * j];
        diana::DevBuf dA(at.data(), at.size() * 8), dQ(at.size() * 8), dR(r.size() * 8);
        diana::check(dt_qr_f64(diana::Device::ctx(), (double *)dQ.p, (double *)dR.p, (double *)dA.p, n, m), "LQ");
        dQ.to_host(q.data(), q.size() * 8);
        dR.to_host(r.data(), r.size() * 8);
        for (size_t i = 0; i < m; i++)
            for (size_t j = 0; j < n; j++) Q[i + m * j] = (Ty)q[j + n * i];
        for (size_t i = 0; i < m; i++)
            for (size_t j = 0; j < m; j++) L[i + m * j] = (Ty)r[j + m * i];
    }
    static void gemm(int tA, int tB, Ty *C, Ty *A, Ty *B, size_t m, size_t n, size_t k) {
        auto a = to_f64(A, m * k), b = to_f64(B, k * n);
        std::vector<double> c(m * n);
        diana::DevBuf dA(a.data(), a.size() * 8), dB(b.data(), b.size() * 8), dC(c.size() * 8);
        const size_t lda = tA ? k : m, ldb = tB ? n : k;
        diana::check(dt_gemm_f64(diana::Device::ctx(), tA, tB, m, n, k, 1.0, (double *)dA.p, lda,
                                 (double *)dB.p, ldb, 0.0, (double *)dC.p, m), "matmul");
        dC.to_host(c.data(), c.size() * 8);
        from_f64(C, c);
    }
    // operator_double.cpp:157-216 (implicit leading dimensions)
    static void matmulNN(Ty *C, Ty *A, Ty *B, size_t m, size_t n, size_t k) { gemm(0, 0, C, A, B, m, n, k); }
    static void matmulNT(Ty *C, Ty *A, Ty *B, size_t m, size_t n, size_t k) { gemm(0, 1, C, A, B, m, n, k); }
    static void matmulTN(Ty *C, Ty *A, Ty *B, size_t m, size_t n, size_t k) { gemm(1, 0, C, A, B, m, n, k); }

    // operator_cpu.tpp:19-29: B (n x m) = A (m x n)^T
    static void transpose(Ty *B, Ty *A, size_t m, size_t n) {
        for (size_t i = 0; i < m; i++)
            for (size_t j = 0; j < n; j++) B[j + i * n] = A[i + j * m];
    }
    // operator_cpu.tpp:56-135: unfolding layout maps (host; the GPU path never
    // materialises them).
    static void tenmat(Ty *B, Ty *A, const shape_t &shape, size_t n) {
        size_t size = Util::calc_size(shape), br = shape[n], bc = 1;
        for (size_t i = 0; i < n; i++) bc *= shape[i];
        for (size_t idx = 0; idx < size; idx += br * bc)
            for (size_t i = 0; i < bc; i++)
                for (size_t j = 0; j < br; j++) B[idx + i * br + j] = A[idx + j * bc + i];
    }
    static void tenmatt(Ty *B, Ty *A, const shape_t &shape, size_t n) {
        size_t size = Util::calc_size(shape), br = shape[n], bc = 1;
        for (size_t i = 0; i < n; i++) bc *= shape[i];
        size_t col = size / br;
        for (size_t idx = 0; idx < size; idx += br * bc)
            for (size_t j = 0; j < br; j++)
                for (size_t i = 0; i < bc; i++) B[j * col + idx / br + i] = A[idx + j * bc + i];
    }
    static void mattten(Ty *B, Ty *A, const shape_t &shape, size_t n) {
        size_t size = Util::calc_size(shape), br = shape[n], bc = 1;
        for (size_t i = 0; i < n; i++) bc *= shape[i];
        size_t col = size / br;
        for (size_t idx = 0; idx < size; idx += br * bc)
            for (size_t j = 0; j < br; j++)
                for (size_t i = 0; i < bc; i++) B[idx + j * bc + i] = A[j * col + idx / br + i];
    }
    // operator_double.cpp:74-82: sqrt of the FP64 sum of squares (on the GPU)
    static double fnorm(Ty *A, size_t n) {
        diana::DevBuf d(A, n * sizeof(Ty));
        double s = 0.0;
        diana::check(dt_sumsq(diana::Device::ctx(), d.p, diana::DType<Ty>::v, n, &s), "fnorm");
        return std::sqrt(s);
    }
    static Ty sum(Ty *A, size_t n) {
        Ty r = 0;
        for (size_t i = 0; i < n; i++) r += A[i];
        return r;
    }
    // operator_double.cpp:218-246: orthonormalise the len2 columns (length len1)
    static void orth1(Ty *A, const size_t len1, const size_t len2) {
        auto a = to_f64(A, len1 * len2);
        diana::DevBuf d(a.data(), a.size() * 8);
        diana::check(dt_qr_f64(diana::Device::ctx(), (double *)d.p, nullptr, (double *)d.p, len1, len2), "orth1");
        d.to_host(a.data(), a.size() * 8);
        from_f64(A, a);
    }
    // operator_double.cpp:248-277: orthonormalise the len1 rows (length len2)
    static void orth2(Ty *A, const size_t len1, const size_t len2) {
        auto a = to_f64(A, len1 * len2);
        diana::DevBuf d(a.data(), a.size() * 8);
        diana::check(dt_orth_rows_f64(diana::Device::ctx(), (double *)d.p, len1, len2), "orth2");
        d.to_host(a.data(), a.size() * 8);
        from_f64(A, a);
    }
    static void AXPY(size_t N, Ty *x, Ty *y, Ty a) { for (size_t i = 0; i < N; i++) y[i] += a * x[i]; }
};
#endif
