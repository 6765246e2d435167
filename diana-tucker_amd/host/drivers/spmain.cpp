// spmain.cpp's workload on one GPU: read a .tns file, build the SpTensor,
// Sp_HOOI_ALS with truncation (default 10) for every mode, one sweep, then the
// Summary table.
//   ./spmain <file.tns> [nthreads (ignored)] [truncation] [sweeps]
#include <cstdio>
#include <cstdlib>
#include "algorithm.hpp"
#include "read_data.hpp"
#include "sp_tensor.hpp"
#include "summary.hpp"

int main(int argc, char *argv[]) {
    if (argc < 2) {
        printf("Usage: ./sptucker [filepath]\n");
        exit(1);
    }
    Summary::init();
    srand((unsigned int)1);
    auto dbf = (data_buffer<double> *)malloc(sizeof(data_buffer<double>));
    read_data<double>(argv[1], dbf);
    size_t truncation = 10;
    if (argc >= 4) truncation = (size_t)strtol(argv[3], nullptr, 10);
    size_t sweeps = 1;
    if (argc >= 5) sweeps = (size_t)strtol(argv[4], nullptr, 10);
    output(std::string("File path  : ") + argv[1]);
    output("truncation : " + std::to_string(truncation));
    SpTensor<double> A(dbf);
    for (size_t n = 0; n < dbf->ndim; n++) free(dbf->index_lists[n]);
    free(dbf->index_lists);
    free(dbf->vals);
    free(dbf->dims);
    free(dbf);
    shape_t R;
    for (size_t i = 0; i < A.ndim(); i++) R.push_back(truncation);
    auto [G, U] = Algorithm::SpTucker::Sp_HOOI_ALS(A, R, sweeps);
    (void)G;
    (void)U;
    Summary::finalize();
    Summary::print_summary();
    return 0;
}
