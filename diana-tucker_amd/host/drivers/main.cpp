// Dense driver with the shape of the reference's main.cpp (main.cpp:13-29):
// srand(20000905), a Gaussian tensor filled by Tensor::randn, HOOI_ALS with
// 5 sweeps, the Summary table at the end.  One rank: partition {1,1,1}.
// Optional arguments: I0 I1 I2 R.
#include <cstdlib>
#include "algorithm.hpp"
#include "summary.hpp"

int main(int argc, char **argv) {
    mpi_init();
    Summary::init();
    srand((unsigned int)20000905);
    shape_t I{500, 400, 300};
    size_t r = 100;
    if (argc >= 5) {
        I = {(size_t)atol(argv[1]), (size_t)atol(argv[2]), (size_t)atol(argv[3])};
        r = (size_t)atol(argv[4]);
    }
    shape_t R{r, r, r};
    shape_t par{1, 1, 1};
    auto *distribution = new DistributionCartesianBlock(par, mpi_rank());
    auto T = Tensor<double>(distribution, I);
    T.randn();
    auto [G, U] = Algorithm::Tucker::HOOI_ALS(T, R, 5);
    (void)G;
    (void)U;
    Summary::finalize();
    Summary::print_summary();
    mpi_finalize();
    return 0;
}
