// The reference's own gtests (tests/testcases/function/distributed/ttm.cpp,
// gram.cpp, fixture FunctionDistributedTest.cpp:8-15) rewritten against this
// build's host layer on one rank: the 5x4x3 fixture is assembled from the
// 2x3x1-grid rule t[i] = 10*rank + i on ceil-div blocks, then TTM1 and Gram1
// are checked element by element (EXPECT_DOUBLE_EQ -> exact equality), plus a
// HOOI run through the same API.  Exit status 0 = all passed.
#define DIANA_THROW_ON_ERROR
#include <cmath>
#include <cstdio>
#include "algorithm.hpp"

static int failures = 0;
#define EXPECT_EQ_D(a, b)                                                                 \
    do {                                                                                  \
        if (!((a) == (b))) {                                                              \
            std::fprintf(stderr, "FAIL %s:%d %s=%.17g != %.17g\n", __FILE__, __LINE__, #a, \
                         (double)(a), (double)(b));                                       \
            failures++;                                                                   \
        }                                                                                 \
    } while (0)

static Tensor<double> fixture(DistributionCartesianBlock *dist) {
    const size_t shape[3] = {5, 4, 3}, par[3] = {2, 3, 1};
    Tensor<double> t(dist, {5, 4, 3});
    for (int rank = 0; rank < 6; rank++) {
        size_t c[3] = {(size_t)rank % par[0], ((size_t)rank / par[0]) % par[1], 0};
        size_t lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = DIANA_CEILDIV(shape[d] * c[d], par[d]);
            hi[d] = DIANA_CEILDIV(shape[d] * (c[d] + 1), par[d]);
        }
        size_t i = 0;
        for (size_t z = lo[2]; z < hi[2]; z++)
            for (size_t y = lo[1]; y < hi[1]; y++)
                for (size_t x = lo[0]; x < hi[0]; x++) t(x, y, z) = 10.0 * rank + (double)(i++);
    }
    return t;
}

int main() {
    mpi_init();
    auto *dist = new DistributionCartesianBlock({1, 1, 1}, mpi_rank());
    auto *dis_global = new DistributionGlobal();
    {  // TTM1 (ttm.cpp:7-37)
        Tensor<double> t = fixture(dist);
        Tensor<double> m1(dis_global, {4, 4});
        for (size_t i = 0; i < m1.size(); i++) m1[i] = (double)i;
        Tensor<double> m2(dis_global, {2, 3});
        for (size_t i = 0; i < m2.size(); i++) m2[i] = (double)i;
        t = Function::ttm<double>(t, m1, 1);
        t = Function::ttm<double>(t, m2, 2);
        auto ans = Function::gather(t);
        double ground_truth[] = {4752.0, 4896.0, 5040.0, 5888.0, 6032.0, 5310.0, 5478.0, 5646.0,
                                 6620.0, 6788.0, 5868.0, 6060.0, 6252.0, 7352.0, 7544.0, 6426.0,
                                 6642.0, 6858.0, 8084.0, 8300.0, 6960.0, 7176.0, 7392.0, 8720.0,
                                 8936.0, 7761.0, 8013.0, 8265.0, 9794.0, 10046.0, 8562.0, 8850.0,
                                 9138.0, 10868.0, 11156.0, 9363.0, 9687.0, 10011.0, 11942.0, 12266.0};
        for (size_t i = 0; i < ans.size(); i++) EXPECT_EQ_D(ans[i], ground_truth[i]);
    }
    {  // Gram1 (gram.cpp:7-19)
        Tensor<double> t = fixture(dist);
        auto ans = Function::gram<double>(t, 1);
        double ground_truth[] = {1990.0, 2353.0, 4487.0, 7487.0, 2353.0, 2821.0, 5525.0, 9305.0,
                                 4487.0, 5525.0, 11599.0, 19819.0, 7487.0, 9305.0, 19819.0, 34039.0};
        for (size_t i = 0; i < ans.size(); i++) EXPECT_EQ_D(ans[i], ground_truth[i]);
    }
    {  // ttt_except, matmul, inverse, QR through the same API
        Tensor<double> t = fixture(dist);
        auto c = Function::ttt_except<double>(t, t, 1);
        auto g = Function::gram<double>(t, 1);
        for (size_t i = 0; i < g.size(); i++) EXPECT_EQ_D(c[i], g[i]);
        Tensor<double> a({3, 3}, false);
        double av[9] = {4, 1, 0, 1, 3, 1, 0, 1, 2};
        for (int i = 0; i < 9; i++) a[i] = av[i];
        auto ai = Function::inverse<double>(a);
        auto e = Function::matmulNN<double>(a, ai);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                if (std::fabs(e[i + 3 * j] - (i == j ? 1.0 : 0.0)) > 1e-12) failures++;
        auto [q, r] = Function::reduced_QR<double>(a);
        auto qr = Function::matmulNN<double>(q, r);
        for (int i = 0; i < 9; i++)
            if (std::fabs(qr[i] - a[i]) > 1e-12) failures++;
    }
    {  // HOOI_ALS: a noise-free rank-(2,2,2) tensor is recovered exactly
        srand(20000905);
        shape_t I{12, 10, 8}, R{2, 2, 2};
        Tensor<double> X(dist, I);
        Tensor<double> g(dis_global, R), u0(dis_global, {12, 2}), u1(dis_global, {10, 2}),
            u2(dis_global, {8, 2});
        g.randn();
        u0.randn();
        u1.randn();
        u2.randn();
        for (size_t k = 0; k < 8; k++)
            for (size_t j = 0; j < 10; j++)
                for (size_t i = 0; i < 12; i++) {
                    double s = 0;
                    for (size_t c = 0; c < 2; c++)
                        for (size_t b = 0; b < 2; b++)
                            for (size_t a = 0; a < 2; a++)
                                s += g(a, b, c) * u0(i, a) * u1(j, b) * u2(k, c);
                    X(i, j, k) = s;
                }
        diana::hooi_options().verbose = false;
        diana::hooi_options().fp64 = true;
        auto [G, U] = Algorithm::Tucker::HOOI_ALS(X, R, 3);
        double res = diana::hooi_options().residuals.back();
        if (!(res < 1e-6)) {
            std::fprintf(stderr, "FAIL HOOI residual %.3e\n", res);
            failures++;
        }
        if (std::fabs(G.fnorm() - X.fnorm()) > 1e-8 * X.fnorm()) failures++;
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
