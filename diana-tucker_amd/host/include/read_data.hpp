// data_buffer and read_data (include/read_data.hpp, src/read_data.cpp:36-129):
// a FROSTT-style .tns text file, one nonzero per line: ndim 1-based unsigned
// indices then the value, separated by single spaces.  ndim = number of
// spaces on the first line; dims[n] = largest 0-based index + 1; a missing
// file prints "Error: Can not find file: <name>" and exits with status 1.
// Parsed on the host (one rank reads everything).
#ifndef DIANA_AMD_READ_DATA_HPP
#define DIANA_AMD_READ_DATA_HPP
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "def.hpp"

template <typename Ty>
class data_buffer {
public:
    size_t **index_lists;
    size_t ndim;
    size_t nnz;
    Ty *vals;
    size_t *dims;
};

namespace diana {
// unsigned decimal at p; advances p past the number and one separator
inline size_t tns_uint(const char *&p) {
    size_t v = 0;
    while (*p >= '0' && *p <= '9') v = v * 10 + (size_t)(*p++ - '0');
    ++p;
    return v;
}
// non-negative decimal (digits and one '.') at p, same advance rule
inline double tns_value(const char *&p, const char *end) {
    const char *s = p;
    while (p < end && ((*p >= '0' && *p <= '9') || *p == '.')) ++p;
    const double v = std::strtod(std::string(s, p).c_str(), nullptr);
    ++p;
    return v;
}
}  // namespace diana

template <typename Ty>
void read_data(char *filename, data_buffer<Ty> *dbf) {
    const int fd = open(filename, O_RDONLY);
    if (fd < 0) {
        std::fprintf(stderr, "Error: Can not find file: %s\n", filename);
        std::exit(1);
    }
    struct stat st;
    fstat(fd, &st);
    const size_t size = (size_t)st.st_size;
    const char *buf = size ? (const char *)mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0) : nullptr;
    if (size && buf == MAP_FAILED) std::exit(1);
    const char *end = buf + size;
    size_t ndim = 0, nnz = 0;
    for (const char *p = buf; p < end && *p != '\n'; ++p)
        if (*p == ' ') ndim++;
    for (const char *p = buf; p < end; ++p)
        if (*p == '\n') nnz++;
    dbf->ndim = ndim;
    dbf->nnz = nnz;
    dbf->index_lists = (size_t **)std::malloc(sizeof(size_t *) * ndim);
    for (size_t n = 0; n < ndim; n++) dbf->index_lists[n] = (size_t *)std::malloc(sizeof(size_t) * (nnz ? nnz : 1));
    dbf->vals = (Ty *)std::malloc(sizeof(Ty) * (nnz ? nnz : 1));
    dbf->dims = (size_t *)std::malloc(sizeof(size_t) * (ndim ? ndim : 1));
    for (size_t n = 0; n < ndim; n++) dbf->dims[n] = 0;
    const char *p = buf;
    for (size_t k = 0; k < nnz; k++) {
        for (size_t n = 0; n < ndim; n++) {
            const size_t i = diana::tns_uint(p) - 1;
            dbf->index_lists[n][k] = i;
            if (dbf->dims[n] < i + 1) dbf->dims[n] = i + 1;
        }
        dbf->vals[k] = (Ty)diana::tns_value(p, end);
    }
    if (buf) munmap((void *)buf, size);
    close(fd);
}
#endif
