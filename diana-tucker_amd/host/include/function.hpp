// Function:: templates of the reference (include/function.hpp:6-94) on one
// rank: the tensor operations run in libdiana_hip.so on the GPU (no unfolding
// copies), the matrix wrappers call Operator<Ty> (also on the GPU).
#ifndef DIANA_AMD_FUNCTION_HPP
#define DIANA_AMD_FUNCTION_HPP
#include <tuple>
#include <vector>
#include "summary.hpp"
#include "tensor.hpp"

namespace Function {
namespace detail {
inline int work_type(int dt) { return dt == DT_F64 ? DT_F64 : DT_F32; }
}

// matrix.tpp:6-53
template <typename Ty>
Tensor<Ty> matmulNN(const Tensor<Ty> &A, const Tensor<Ty> &B) {
    if (!A.is_matrix() || A.shape()[1] != B.shape()[0]) error("matmulNN: shape mismatch");
    size_t m = A.shape()[0], n = B.shape()[1], k = A.shape()[1];
    SummaryScope sc("Function::matmulNN()", 2LL * m * n * k, (long long)(m * k + k * n + m * n) * sizeof(Ty));
    Tensor<Ty> ret({m, n}, false);
    A.op()->matmulNN(ret.data(), A.data(), B.data(), m, n, k);
    return ret;
}
template <typename Ty>
Tensor<Ty> matmulNT(const Tensor<Ty> &A, const Tensor<Ty> &B) {
    if (!A.is_matrix() || A.shape()[1] != B.shape()[1]) error("matmulNT: shape mismatch");
    size_t m = A.shape()[0], n = B.shape()[0], k = A.shape()[1];
    SummaryScope sc("Function::matmulNT()", 2LL * m * n * k, (long long)(m * k + k * n + m * n) * sizeof(Ty));
    Tensor<Ty> ret({m, n}, false);
    A.op()->matmulNT(ret.data(), A.data(), B.data(), m, n, k);
    return ret;
}
template <typename Ty>
Tensor<Ty> matmulTN(const Tensor<Ty> &A, const Tensor<Ty> &B) {
    if (!A.is_matrix() || A.shape()[0] != B.shape()[0]) error("matmulTN: shape mismatch");
    size_t m = A.shape()[1], n = B.shape()[1], k = A.shape()[0];
    SummaryScope sc("Function::matmulTN()", 2LL * m * n * k, (long long)(m * k + k * n + m * n) * sizeof(Ty));
    Tensor<Ty> ret({m, n}, false);
    A.op()->matmulTN(ret.data(), A.data(), B.data(), m, n, k);
    return ret;
}
// matrix.tpp:55-66
template <typename Ty>
Tensor<Ty> inverse(const Tensor<Ty> &A) {
    if (!A.is_matrix() || A.shape()[0] != A.shape()[1]) error("inverse: not square");
    Tensor<Ty> ret(A.shape(), true);
    A.op()->inverse(ret.data(), A.data(), A.shape()[0]);
    return ret;
}
// matrix.tpp:68-100
template <typename Ty>
std::tuple<Tensor<Ty>, Tensor<Ty>> reduced_LQ(const Tensor<Ty> &A) {
    if (!A.is_matrix() || A.shape()[0] > A.shape()[1]) error("reduced_LQ: need m <= n");
    size_t m = A.shape()[0], n = A.shape()[1];
    Tensor<Ty> L({m, m}, true), Q({m, n}, true);
    A.op()->LQ(L.data(), Q.data(), A.data(), m, n);
    return std::make_tuple(L, Q);
}
template <typename Ty>
std::tuple<Tensor<Ty>, Tensor<Ty>> reduced_QR(const Tensor<Ty> &A) {
    if (!A.is_matrix() || A.shape()[0] < A.shape()[1]) error("reduced_QR: need m >= n");
    size_t m = A.shape()[0], n = A.shape()[1];
    Tensor<Ty> Q({m, n}, true), R({n, n}, true);
    A.op()->QR(Q.data(), R.data(), A.data(), m, n);
    return std::make_tuple(Q, R);
}
// matrix.tpp:102-114
template <typename Ty>
Tensor<Ty> transpose(const Tensor<Ty> &A) {
    size_t m = A.shape()[0], n = A.shape()[1];
    Tensor<Ty> At({n, m}, false);
    A.op()->transpose(At.data(), A.data(), m, n);
    return At;
}
// matrix.tpp:117-129
template <typename Ty>
Tensor<Ty> gram(const Tensor<Ty> &A) {
    size_t M = A.shape()[0], N = A.shape()[1];
    Tensor<Ty> ret({M, M}, false);
    A.op()->matmulNT(ret.data(), A.data(), A.data(), M, M, N);
    return ret;
}

// tensor.tpp:272-335: Y = A x_n M, M is J x I_n.
template <typename Ty>
Tensor<Ty> ttm(const Tensor<Ty> &A, const Tensor<Ty> &M, size_t n) {
    if (!M.is_matrix() || A.shape()[n] != M.shape()[1]) error("ttm: M must be J x I_n");
    const int dt = diana::DType<Ty>::v, wt = detail::work_type(dt);
    shape_t out = A.shape();
    out[n] = M.shape()[0];
    // 2 |A| J flop; A and Y each moved once
    SummaryScope sc("Function::ttm()", 2LL * (long long)A.size() * (long long)M.shape()[0],
                    (long long)(A.size() + A.size() / A.shape()[n] * M.shape()[0]) * (long long)sizeof(Ty));
    Tensor<Ty> ret(A.distribution(), out, false);
    diana::DevBuf dA(A.data(), A.size() * sizeof(Ty)), dM(M.data(), M.size() * sizeof(Ty));
    diana::DevBuf dY(ret.size() * (wt == DT_F64 ? 8 : 4));
    diana::check(dt_ttm(diana::Device::ctx(), dA.p, dt, (int)A.ndim(), A.shape().data(), dM.p, dt,
                        M.shape()[0], (int)n, dY.p, wt), "ttm");
    dY.to_host(ret.data(), ret.size() * sizeof(Ty));
    return ret;
}

// tensor.tpp:337-348
template <typename Ty>
Tensor<Ty> ttmc(const Tensor<Ty> &A, const std::vector<Tensor<Ty>> &M, const std::vector<size_t> &idx) {
    if (M.size() != idx.size()) error("ttmc: |M| != |idx|");
    Tensor<Ty> r = A;
    for (size_t i = 0; i < M.size(); i++) r = ttm(r, M[i], idx[i]);
    return r;
}

// tensor.tpp:18-120: A_(n) A_(n)^T
template <typename Ty>
Tensor<Ty> gram(const Tensor<Ty> &A, size_t n) {
    const int dt = diana::DType<Ty>::v;
    const size_t I = A.shape()[n];
    // symmetric rank-K update, K = |A| / I: I (I + 1) K flop, A read once
    SummaryScope sc("Function::gram()", (long long)(I + 1) * (long long)A.size(),
                    (long long)A.size() * (long long)sizeof(Ty));
    Tensor<Ty> ret({I, I}, false);
    std::vector<double> s(I * I);
    diana::DevBuf dA(A.data(), A.size() * sizeof(Ty)), dS(s.size() * 8);
    diana::check(dt_gram(diana::Device::ctx(), dA.p, dt, (int)A.ndim(), A.shape().data(), (int)n, dS.p,
                         DT_F64), "gram");
    dS.to_host(s.data(), s.size() * 8);
    for (size_t i = 0; i < s.size(); i++) ret[i] = (Ty)s[i];
    return ret;
}

// tensor.tpp:132-259: A_(n) B_(n)^T
template <typename Ty>
Tensor<Ty> ttt_except(const Tensor<Ty> &A, const Tensor<Ty> &B, size_t n) {
    for (size_t i = 0; i < A.ndim(); i++)
        if (i != n && A.shape()[i] != B.shape()[i]) error("ttt_except: shapes differ off mode n");
    const size_t I = A.shape()[n], J = B.shape()[n];
    SummaryScope sc("Function::ttt_except()", 2LL * (long long)A.size() * (long long)J,
                    (long long)(A.size() + B.size()) * (long long)sizeof(Ty));
    Tensor<Ty> ret({I, J}, false);
    std::vector<double> c(I * J);
    diana::DevBuf dA(A.data(), A.size() * sizeof(Ty)), dB(B.data(), B.size() * sizeof(Ty)), dC(c.size() * 8);
    diana::check(dt_ttt_except(diana::Device::ctx(), dA.p, dB.p, diana::DType<Ty>::v, (int)A.ndim(),
                               A.shape().data(), J, (int)n, (double *)dC.p), "ttt_except");
    dC.to_host(c.data(), c.size() * 8);
    for (size_t i = 0; i < c.size(); i++) ret[i] = (Ty)c[i];
    return ret;
}

// tensor.tpp:467-480
template <typename Ty>
double fnorm(const Tensor<Ty> &A) { return A.fnorm(); }
template <typename Ty>
Ty sum(const Tensor<Ty> &A) { return A.sum(); }

// tensor.tpp:350-465 at one rank: the block already is the global tensor.
template <typename Ty>
Tensor<Ty> gather(const Tensor<Ty> &A) { return A.copy(); }
}  // namespace Function
#endif
