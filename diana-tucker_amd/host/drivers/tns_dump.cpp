// Prints what read_data parses from a .tns file (host only, no GPU):
// "ndim nnz", "dims d0 d1 ...", then one line per nonzero "i0 i1 ... value".
#include <cstdio>
#include "read_data.hpp"

int main(int argc, char *argv[]) {
    if (argc < 2) return 2;
    data_buffer<double> dbf;
    read_data<double>(argv[1], &dbf);
    printf("%zu %zu\ndims", dbf.ndim, dbf.nnz);
    for (size_t n = 0; n < dbf.ndim; n++) printf(" %zu", dbf.dims[n]);
    printf("\n");
    for (size_t k = 0; k < dbf.nnz; k++) {
        for (size_t n = 0; n < dbf.ndim; n++) printf("%zu ", dbf.index_lists[n][k]);
        printf("%.17g\n", dbf.vals[k]);
    }
    return 0;
}
