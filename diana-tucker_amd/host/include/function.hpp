// Function:: templates of the reference (include/function.hpp:6-94) on one
// rank: the tensor operations run in libdiana_hip.so on the GPU (no unfolding
// copies), the matrix wrappers call Operator<Ty> (also on the GPU).
#ifndef DIANA_AMD_FUNCTION_HPP
#define DIANA_AMD_FUNCTION_HPP
#include <tuple>
#include <vector>
#include "sp_tensor.hpp"
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

// ---- sparse (include/function.hpp:70-90, template/function/sp_tensor.tpp) ----
namespace detail {
// device copies of the transposed factors Ut[m] (R_m x I_m column-major FP64),
// mode n's taken from `Rn` (the reference passes it separately)
template <typename Ty>
struct SpFactors {
    std::vector<diana::DevBuf *> buf;
    std::vector<const double *> ptr;
    std::vector<size_t> ranks;
    SpFactors(const SpTensor<Ty> &A, const std::vector<Tensor<Ty>> &M, size_t n, const Tensor<Ty> &Rn) {
        if (M.size() != A.ndim()) error("sparse TTMc: one factor per mode");
        for (size_t m = 0; m < A.ndim(); m++) {
            const Tensor<Ty> &U = m == n ? Rn : M[m];
            if (!U.is_matrix() || U.shape()[1] != A.shape()[m]) error("sparse TTMc: factor " + std::to_string(m) + " must be R x I");
            std::vector<double> h(U.data(), U.data() + U.size());
            buf.push_back(new diana::DevBuf(h.data(), h.size() * 8));
            ptr.push_back((const double *)buf.back()->p);
            ranks.push_back(U.shape()[0]);
        }
    }
    ~SpFactors() {
        for (auto *b : buf) delete b;
    }
    size_t K(size_t n) const {
        size_t k = 1;
        for (size_t m = 0; m < ranks.size(); m++)
            if (m != n) k *= ranks[m];
        return k;
    }
};
}  // namespace detail

// sp_tensor.tpp:236-286 as intended: L = Y_(n)ᵀ U_n (prod_{m != n} R_m x R_n),
// Y_(n) the slices' Kronecker rows of the other modes' factors (first other
// mode fastest), U_n = Rᵀ.  M[m]: Ut[m], R_m x I_m.
template <typename Ty>
Tensor<Ty> ttmc_mTR(const SpTensor<Ty> &A, const std::vector<Tensor<Ty>> &M, size_t n, const Tensor<Ty> &R) {
    detail::SpFactors<Ty> f(A, M, n, R);
    const size_t K = f.K(n), Rn = f.ranks[n];
    SummaryScope sc("Function::ttmc_mTR()", (long long)A.nnz() * (long long)(2 * K + Rn),
                    (long long)A.nnz() * (long long)(4 * A.ndim() + (A.fp64() ? 8 : 4)));
    diana::DevBuf dL(K * Rn * 8);
    diana::check(dt_sp_ttmc_mtr(diana::Device::ctx(), A.device(), (int)n, f.ptr.data(), f.ranks.data(),
                                (double *)dL.p), "ttmc_mTR");
    std::vector<double> h(K * Rn);
    dL.to_host(h.data(), h.size() * 8);
    static DistributionGlobal global;
    Tensor<Ty> ret(&global, {K, Rn}, false);
    for (size_t i = 0; i < h.size(); i++) ret[i] = (Ty)h[i];
    return ret;
}

// sp_tensor.tpp:288-365 as intended: ret (R_n x I_n) = (Y_(n) L)ᵀ; slices
// without nonzeros give zero columns.
template <typename Ty>
void ttmc_mTL(const SpTensor<Ty> &A, const std::vector<Tensor<Ty>> &M, size_t n, const Tensor<Ty> &L, Tensor<Ty> &ret) {
    detail::SpFactors<Ty> f(A, M, n, M[n]);
    const size_t K = f.K(n), Rn = L.shape()[1];
    if (!L.is_matrix() || L.shape()[0] != K || Rn != f.ranks[n]) error("ttmc_mTL: L must be prod R_{-n} x R_n");
    if (ret.size() != Rn * A.shape()[n]) error("ttmc_mTL: ret must be R_n x I_n");
    SummaryScope sc("Function::ttmc_mTL()", (long long)A.nnz() * (long long)(2 * K) + 2LL * (long long)(A.shape()[n] * K * Rn),
                    (long long)A.nnz() * (long long)(4 * A.ndim() + (A.fp64() ? 8 : 4)));
    std::vector<double> hl(L.data(), L.data() + L.size());
    diana::DevBuf dL(hl.data(), hl.size() * 8), dR(ret.size() * 8);
    diana::check(dt_sp_ttmc_mtl(diana::Device::ctx(), A.device(), (int)n, f.ptr.data(), f.ranks.data(),
                                (const double *)dL.p, (double *)dR.p), "ttmc_mTL");
    std::vector<double> h(ret.size());
    dR.to_host(h.data(), h.size() * 8);
    for (size_t i = 0; i < h.size(); i++) ret[i] = (Ty)h[i];
}

// sp_tensor.tpp:6-14: sqrt of the FP64 sum of squares of the values
template <typename Ty>
double fnorm(const SpTensor<Ty> &A) {
    int sm = 0;
    int64_t nnz = 0;
    double nrm = 0.0;
    diana::check(dt_sp_info(A.device(), &sm, &nnz, &nrm), "fnorm");
    return nrm;
}

// tensor.tpp:350-465 at one rank: the block already is the global tensor.
template <typename Ty>
Tensor<Ty> gather(const Tensor<Ty> &A) { return A.copy(); }
}  // namespace Function
#endif
