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
    bool fp64 = true;                 // Tensor<double>: compute in FP64 as the reference does;
                                      // false narrows A to FP32 and runs the FP32 kernels
    bool verbose = true;
    std::vector<double> residuals;    // filled by HOOI_ALS
};
inline HooiOptions &hooi_options() {
    static HooiOptions o;
    return o;
}

// Algorithmic work of one dense HOOI sweep for the Summary table: 2 in J flop
// per mode product (in = elements read, J = the rank it contracts to) and
// the elements each product reads and writes, A in its own type and the
// intermediates in the working type (FP32, FP64 for an FP64 tensor).  The
// reference schedule runs N + 1 chains from A (hooi_als.tpp:94-109); the
// dimension tree (sweep_once in dt_hooi.hip) runs two, whose first two
// products are one pass (their intermediate is never stored).  The ALS
// updates (R x R and I x R work) are left out.
struct SweepWork {
    long long flop = 0, bytes = 0;
};
inline SweepWork hooi_sweep_work(const shape_t &I, const shape_t &R, int schedule, size_t a_bytes) {
    const size_t N = I.size(), hm = N / 2, wb = a_bytes == 8 ? 8 : 4;
    SweepWork w;
    auto chain = [&](shape_t sh, const std::vector<size_t> &modes, bool from_A, bool fused) {
        for (size_t k = 0; k < modes.size(); k++) {
            const size_t m = modes[k];
            double in = 1.0;
            for (size_t x : sh) in *= (double)x;
            const double out = in / (double)sh[m] * (double)R[m];
            w.flop += (long long)(2.0 * in * (double)R[m]);
            if (!(fused && k == 1)) w.bytes += (long long)(in * (double)((from_A && k == 0) ? a_bytes : wb));
            if (!(fused && k == 0)) w.bytes += (long long)(out * (double)wb);
            sh[m] = R[m];
        }
        return sh;
    };
    if (schedule == DT_SCHED_REFERENCE || N < 3) {
        for (size_t n = 0; n < N; n++) {
            std::vector<size_t> ms;
            for (size_t i = 0; i < N; i++)
                if (i != n) ms.push_back(i);
            chain(I, ms, true, false);
        }
        std::vector<size_t> all;
        for (size_t i = 0; i < N; i++) all.push_back(i);
        chain(I, all, true, false);
        return w;
    }
    std::vector<size_t> m1{N - 2, N - 1};
    for (size_t i = N - 3; i + 1 > hm && i < N; i--) m1.push_back(i);
    const shape_t TL = chain(I, m1, true, true);
    for (size_t n = 0; n < hm; n++) {
        std::vector<size_t> ms;
        for (size_t i = hm; i-- > 0;)
            if (i != n) ms.push_back(i);
        chain(TL, ms, false, false);
    }
    std::vector<size_t> m2;
    for (size_t i = hm; i-- > 0;) m2.push_back(i);
    const shape_t TR = chain(I, m2, true, m2.size() >= 2);
    shape_t last = TR;
    for (size_t n = hm; n < N; n++) {
        std::vector<size_t> ms;
        for (size_t i = N; i-- > hm;)
            if (i != n) ms.push_back(i);
        last = chain(TR, ms, false, false);
    }
    chain(last, {N - 1}, false, false);
    return w;
}

// One sparse HOOI sweep (mode-sorted copies): per mode the TTMc reads every
// nonzero (N indices and a value), writes Y_(n) and reads the other factors
// once, nnz (2 prod_{m != n} R_m + R_n) flop (the Kronecker rows and the
// Y_(n)ᵀU_n product).
inline SweepWork sp_sweep_work(const shape_t &I, const shape_t &R, long long nnz) {
    SweepWork w;
    for (size_t n = 0; n < I.size(); n++) {
        long long kr = 1, fac = 0;
        for (size_t m = 0; m < I.size(); m++)
            if (m != n) {
                kr *= (long long)R[m];
                fac += (long long)(I[m] * R[m]) * 4;
            }
        w.flop += nnz * (2 * kr + (long long)R[n]);
        w.bytes += nnz * (4 * (long long)I.size() + 4) + (long long)I[n] * kr * 4 + fac;
    }
    return w;
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
    SummaryScope region("Tucker::HOOI_ALS()");
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
    const diana::SweepWork sw = diana::hooi_sweep_work(I, R, opt.schedule, f64 ? 8 : 4);
    for (size_t it = 0; it < max_iter && st == DT_OK; it++) {
        if (opt.verbose) output("Calculating iteration " + std::to_string(it + 1) + " ...");
        double r = 0.0;
        {
            SummaryScope sc("Tucker::HOOI_ALS::sweep()", sw.flop, sw.bytes);
            st = dt_hooi_sweep(h, 0, &r);
        }
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
// last mode, with the reference's output lines: "||A||_F", per sweep
// "Calculating iteration k ..." and each mode's ||L||_F (printed as ||G||_F
// by Sp_TTMC_ALS_, :51-53), then the core's norm and the residual.  Returns
// (G, Ut) with Ut[n] stored as {R_n, I_n} like the reference.  Any number of
// modes; ranks <= 64 with prod_{m != n} R_m <= 4096 (3-way FP32 tensors with
// ranks <= 16 take the MFMA pipeline).
template <typename Ty>
std::tuple<Tensor<Ty>, std::vector<Tensor<Ty>>> Sp_HOOI_ALS(const SpTensor<Ty> &A, const shape_t &R,
                                                            size_t max_iter) {
    const size_t N = A.ndim();
    if (R.size() != N) error("Sp_HOOI_ALS: |R| != ndim");
    SummaryScope region("SpTucker::Sp_HOOI_ALS()");
    const shape_t &I = A.shape();
    dt_ctx *ctx = diana::Device::ctx();
    const double a_norm = Function::fnorm(A);
    output("||A||_F = " + std::to_string(a_norm));
    static DistributionGlobal global;
    std::vector<diana::DevBuf *> dU0, dU;
    std::vector<const double *> u0p;
    std::vector<double *> up;
    for (size_t n = 0; n < N; n++) {
        Tensor<double> Ut(&global, {R[n], I[n]}, false);
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
    dt_sp_hooi *h = nullptr;
    diana::check(dt_sp_hooi_create(ctx, A.device(), R.data(), 0, u0p.data(), &h), "Sp_HOOI_ALS");
    int64_t nnz = 0;
    {
        int sm = 0;
        double nrm = 0.0;
        dt_sp_info(A.device(), &sm, &nnz, &nrm);
    }
    const diana::SweepWork sw = diana::sp_sweep_work(I, R, (long long)nnz);
    std::vector<double> ln(N);
    int st = DT_OK;
    for (size_t it = 0; it < max_iter && st == DT_OK; it++) {
        output("Calculating iteration " + std::to_string(it + 1) + " ...");
        double res = 0.0;
        {
            SummaryScope sc("SpTucker::Sp_HOOI_ALS::sweep()", sw.flop, sw.bytes);
            st = dt_sp_hooi_sweep(h, ln.data(), &res);
        }
        if (st == DT_OK)
            for (size_t n = 0; n < N; n++) output("||G||_F = " + std::to_string(ln[n]));
    }
    if (st == DT_OK) st = dt_sp_hooi_get(h, up.data(), max_iter > 0 ? (double *)dG.p : nullptr, nullptr);
    const std::string err = st == DT_OK ? "" : dt_last_error(ctx);
    dt_sp_hooi_destroy(h);
    if (st != DT_OK) error("Sp_HOOI_ALS: " + err + " (status " + std::to_string(st) + ")");
    std::vector<Tensor<Ty>> Ut;
    std::vector<double> tmp;
    for (size_t n = 0; n < N; n++) {
        tmp.resize(I[n] * R[n]);
        dU[n]->to_host(tmp.data(), tmp.size() * 8);
        Tensor<Ty> ut(&global, {R[n], I[n]}, false);
        for (size_t i = 0; i < I[n]; i++)
            for (size_t r = 0; r < R[n]; r++) ut[r + R[n] * i] = (Ty)tmp[i + I[n] * r];
        Ut.push_back(ut);
        delete dU0[n];
        delete dU[n];
    }
    Tensor<Ty> G(&global, R, false);
    tmp.assign(csz, 0.0);
    if (max_iter > 0) dG.to_host(tmp.data(), csz * 8);
    double g2 = 0.0;
    for (size_t i = 0; i < csz; i++) {
        G[i] = (Ty)tmp[i];
        g2 += tmp[i] * tmp[i];
    }
    output("||G||_F = " + std::to_string(std::sqrt(g2)));
    output("Residual: sqrt(1 - ||G||_F^2 / ||A||_F^2) = " +
           std::to_string(std::sqrt(std::max(0.0, 1 - g2 / (a_norm * a_norm)))));
    return std::make_tuple(G, Ut);
}
}  // namespace SpTucker
}  // namespace Algorithm
#endif
