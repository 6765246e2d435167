// Algorithm:: entry points of the reference (include/algorithm.hpp:10-24) on
// one GPU.  HOOI_ALS keeps the reference's signature, its random start
// (Tensor::randn -> Util::randn after the caller's srand, then reduced_QR,
// hooi_als.tpp:76-81) and its per-sweep output lines (hooi_als.tpp:85-114);
// the sweeps run in libdiana_hip.so.
#ifndef DIANA_AMD_ALGORITHM_HPP
#define DIANA_AMD_ALGORITHM_HPP
#include <cmath>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>
#include "function.hpp"
#include "sp_tensor.hpp"

namespace diana {
// Knobs the reference does not have; defaults reproduce its behaviour on the GPU.
struct HooiOptions {
    int init = DT_INIT_QR;            // reference: randn + reduced_QR
    int schedule = DT_SCHED_TREE;     // DT_SCHED_REFERENCE = the reference's N+1 chains
    int inner_iter = 5;               // ALS_ default (hooi_als.tpp:9)
    bool fp64 = false;                // Tensor<double>: compute in FP64 instead of FP32
    bool verbose = true;
    std::vector<double> residuals;    // filled by HOOI_ALS
};
inline HooiOptions &hooi_options() {
    static HooiOptions o;
    return o;
}
}  // namespace diana

namespace Algorithm {
namespace Tucker {
template <typename Ty>
std::tuple<Tensor<Ty>, std::vector<Tensor<Ty>>> HOOI_ALS(const Tensor<Ty> &A, const shape_t &R,
                                                         size_t max_iter) {
    auto &opt = diana::hooi_options();
    const size_t N = A.ndim();
    if (R.size() != N) error("HOOI_ALS: |R| != ndim");
    const shape_t &I = A.shape_global();
    if (opt.verbose) output("Start Tucker::HOOI_ALS decomposition.. with max_iter = " + std::to_string(max_iter));
    dt_ctx *ctx = diana::Device::ctx();
    // start factors exactly as the reference draws them
    auto dglobal = new DistributionGlobal();
    std::vector<Tensor<double>> Uraw;
    for (size_t n = 0; n < N; n++) {
        Tensor<double> U_rand(dglobal, {I[n], R[n]}, false);
        if (opt.init == DT_INIT_QR) U_rand.randn();
        Uraw.push_back(U_rand);
    }
    // device copies
    const bool f64 = std::is_same<Ty, double>::value && opt.fp64;
    const int dtA = f64 ? DT_F64 : DT_F32;
    std::vector<float> a32;
    const void *hostA = A.data();
    if (!f64 && std::is_same<Ty, double>::value) {
        a32.assign(A.data(), A.data() + A.size());
        hostA = a32.data();
    }
    diana::DevBuf dA(hostA, A.size() * (f64 ? 8 : 4));
    std::vector<diana::DevBuf *> dU, dUout;
    dt_hooi_args args{};
    args.ndim = (int)N;
    args.dtype = dtA;
    args.A = dA.p;
    for (size_t n = 0; n < N; n++) {
        args.shape[n] = I[n];
        args.ranks[n] = R[n];
        dU.push_back(new diana::DevBuf(Uraw[n].data(), Uraw[n].size() * 8));
        dUout.push_back(new diana::DevBuf(I[n] * R[n] * 8));
        args.U_init[n] = (const double *)dU[n]->p;
    }
    args.max_iter = (int)max_iter;
    args.inner_iter = opt.inner_iter;
    args.init = opt.init;
    args.schedule = opt.schedule;
    // the residual is always formed with FP64 accumulation, as the reference's
    // FP64 fnorm of an FP64 core (hooi_als.tpp:105-114)
    args.flags = DT_FLAG_FP64_FIT;
    dt_hooi *h = nullptr;
    diana::check(dt_hooi_create(ctx, &args, &h), "HOOI_ALS");
    dt_hooi_result res{};
    opt.residuals.clear();
    size_t csz = 1;
    for (auto r : R) csz *= r;
    diana::DevBuf dG(csz * 8);
    int st = DT_OK;
    double a2 = 0.0;
    diana::check(dt_sumsq(ctx, dA.p, dtA, A.size(), &a2), "HOOI_ALS fnorm");
    if (opt.verbose) output("||A||_F = " + std::to_string(std::sqrt(a2)));
    for (size_t it = 0; it < max_iter && st == DT_OK; it++) {
        if (opt.verbose) output("Calculating iteration " + std::to_string(it + 1) + " ...");
        double r = 0.0;
        st = dt_hooi_sweep(h, 0, &r);
        if (st == DT_OK) {
            opt.residuals.push_back(r);
            if (opt.verbose) {
                output("||G||_F = " + std::to_string(std::sqrt(a2 * (1.0 - r * r))));
                output("Residual: sqrt(1 - ||G||_F^2 / ||A||_F^2) = " + std::to_string(r));
            }
        }
    }
    for (size_t n = 0; n < N; n++) res.U[n] = (double *)dUout[n]->p;
    res.G = (double *)dG.p;
    if (st == DT_OK) st = dt_hooi_get(h, &res);
    dt_hooi_destroy(h);
    diana::check(st, "HOOI_ALS");
    std::vector<Tensor<Ty>> U;
    std::vector<double> tmp;
    for (size_t n = 0; n < N; n++) {
        Tensor<Ty> u(dglobal, {I[n], R[n]}, false);
        tmp.resize(u.size());
        dUout[n]->to_host(tmp.data(), tmp.size() * 8);
        for (size_t i = 0; i < tmp.size(); i++) u[i] = (Ty)tmp[i];
        U.push_back(u);
        delete dU[n];
        delete dUout[n];
    }
    Tensor<Ty> G(A.distribution(), R, false);
    tmp.resize(csz);
    dG.to_host(tmp.data(), csz * 8);
    for (size_t i = 0; i < csz; i++) G[i] = (Ty)tmp[i];
    if (opt.verbose) output("Done!");
    return std::make_tuple(G, U);
}
}  // namespace Tucker

namespace SpTucker {
// Sp_HOOI_ALS (include/algorithm.hpp:16-19, sp_hooi_als.tpp:67-118) on one
// GPU: the reference's start (per mode, Tensor {R_n, I_n}.randn() after the
// caller's srand, orthonormalised; the transposed factor U_n = Ut_nᵀ), then
// per sweep and mode U_n <- orth(Y_(n) Y_(n)ᵀ U_n) and the core from the
// last mode, with the reference's output lines.  Returns (G, Ut) with Ut[n]
// stored as {R_n, I_n} like the reference.  Ranks up to 16, 3-way tensors.
template <typename Ty>
std::tuple<Tensor<Ty>, std::vector<Tensor<Ty>>> Sp_HOOI_ALS(const SpTensor<Ty> &A, const shape_t &R,
                                                            size_t max_iter) {
    const size_t N = A.ndim();
    if (R.size() != N) error("Sp_HOOI_ALS: |R| != ndim");
    const shape_t &I = A.shape();
    dt_ctx *ctx = diana::Device::ctx();
    double a2 = 0.0;
    {
        int sm = 0;
        int64_t nnz = 0;
        double nrm = 0.0;
        dt_sp_info(A.device(), &sm, &nnz, &nrm);
        a2 = nrm * nrm;
    }
    output("||A||_F = " + std::to_string(std::sqrt(a2)));
    auto dglobal = new DistributionGlobal();
    std::vector<diana::DevBuf *> dU0, dU;
    std::vector<const double *> u0p;
    std::vector<double *> up;
    for (size_t n = 0; n < N; n++) {
        Tensor<double> Ut(dglobal, {R[n], I[n]}, false);
        Ut.randn();
        std::vector<double> U(I[n] * R[n]);
        for (size_t i = 0; i < I[n]; i++)
            for (size_t r = 0; r < R[n]; r++) U[i + I[n] * r] = Ut[r + R[n] * i];
        dU0.push_back(new diana::DevBuf(U.data(), U.size() * 8));
        dU.push_back(new diana::DevBuf(U.size() * 8));
        u0p.push_back((const double *)dU0.back()->p);
        up.push_back((double *)dU.back()->p);
    }
    size_t csz = 1;
    for (auto r : R) csz *= r;
    diana::DevBuf dG(csz * 8);
    std::vector<double> per(max_iter ? max_iter : 1);
    double res = 0.0, ms = 0.0;
    for (size_t it = 0; it < max_iter; it++) output("Calculating iteration " + std::to_string(it + 1) + " ...");
    diana::check(dt_hooi_sparse(ctx, A.device(), R.data(), (int)max_iter, 0, u0p.data(), up.data(),
                                (double *)dG.p, per.data(), &res, &ms),
                 "Sp_HOOI_ALS");
    std::vector<Tensor<Ty>> Ut;
    std::vector<double> tmp;
    for (size_t n = 0; n < N; n++) {
        tmp.resize(I[n] * R[n]);
        dU[n]->to_host(tmp.data(), tmp.size() * 8);
        Tensor<Ty> ut(dglobal, {R[n], I[n]}, false);
        for (size_t i = 0; i < I[n]; i++)
            for (size_t r = 0; r < R[n]; r++) ut[r + R[n] * i] = (Ty)tmp[i + I[n] * r];
        Ut.push_back(ut);
        delete dU0[n];
        delete dU[n];
    }
    Tensor<Ty> G(dglobal, R, false);
    tmp.resize(csz);
    dG.to_host(tmp.data(), csz * 8);
    double g2 = 0.0;
    for (size_t i = 0; i < csz; i++) {
        G[i] = (Ty)tmp[i];
        g2 += tmp[i] * tmp[i];
    }
    output("||G||_F = " + std::to_string(std::sqrt(g2)));
    output("Residual: sqrt(1 - ||G||_F^2 / ||A||_F^2) = " + std::to_string(std::sqrt(std::max(0.0, 1 - g2 / a2))));
    return std::make_tuple(G, Ut);
}
}  // namespace SpTucker
}  // namespace Algorithm
#endif
