// data_buffer and read_data (include/read_data.hpp, src/read_data.cpp:36-129):
// a FROSTT-style .tns text file, one nonzero per line: ndim 1-based unsigned
// indices then the value.  ndim = the number of tokens on the first line
// minus one; dims[n] = largest 0-based index + 1; a missing file prints
// "Error: Can not find file: <name>" and exits with status 1 (as the
// reference does).  Parsed on the host (one rank reads everything).
//
// Unlike the reference's character scanner (read_data.cpp:13-33: digits and
// '.' only, so a sign or an exponent shifts every later field), each line is
// split into tokens on spaces and tabs (a '\r' before the '\n' is ignored),
// indices are read as unsigned integers and the value with the full
// floating-point syntax (sign, exponent), every read bounded by the line.
// Blank lines and lines starting with '#' are skipped; a line with a wrong
// token count, an index of 0 or an unreadable number is an error naming the
// file and line.
#ifndef DIANA_AMD_READ_DATA_HPP
#define DIANA_AMD_READ_DATA_HPP
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "def.hpp"
#include "logger.hpp"

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
struct TnsToken {
    const char *b, *e;
};
inline bool tns_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }
// tokens of the line [p, end) (end = the '\n' or the end of the file)
inline void tns_tokens(const char *p, const char *end, std::vector<TnsToken> &out) {
    out.clear();
    while (p < end) {
        while (p < end && tns_space(*p)) ++p;
        if (p >= end) break;
        const char *b = p;
        while (p < end && !tns_space(*p)) ++p;
        out.push_back({b, p});
    }
}
inline bool tns_skip(const std::vector<TnsToken> &t) { return t.empty() || *t[0].b == '#'; }
[[noreturn]] inline void tns_fail(const char *file, size_t line, const std::string &what) {
    error(std::string("read_data: ") + file + ":" + std::to_string(line) + ": " + what);
}
inline size_t tns_index(const TnsToken &t, const char *file, size_t line) {
    size_t v = 0;
    const auto r = std::from_chars(t.b, t.e, v);
    if (r.ec != std::errc() || r.ptr != t.e) tns_fail(file, line, "index '" + std::string(t.b, t.e) + "' is not an unsigned integer");
    if (v == 0) tns_fail(file, line, "index 0 (indices are 1-based)");
    return v;
}
inline double tns_value(const TnsToken &t, const char *file, size_t line) {
    const char *b = t.b;
    if (b < t.e && *b == '+') ++b;  // from_chars takes no leading '+'
    double v = 0.0;
    const auto r = std::from_chars(b, t.e, v);
    if (r.ec != std::errc() || r.ptr != t.e) tns_fail(file, line, "value '" + std::string(t.b, t.e) + "' is not a number");
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
    std::vector<diana::TnsToken> tok;
    // first pass: ndim from the first data line, count the data lines
    size_t ndim = 0, nnz = 0;
    bool have_ndim = false;
    for (const char *p = buf; p < end;) {
        const char *e = p;
        while (e < end && *e != '\n') ++e;
        diana::tns_tokens(p, e, tok);
        if (!diana::tns_skip(tok)) {
            if (!have_ndim) {
                if (tok.size() < 2) error(std::string("read_data: ") + filename + ": first line needs indices and a value");
                ndim = tok.size() - 1;
                have_ndim = true;
            }
            nnz++;
        }
        p = e < end ? e + 1 : end;
    }
    dbf->ndim = ndim;
    dbf->nnz = nnz;
    dbf->index_lists = (size_t **)std::malloc(sizeof(size_t *) * (ndim ? ndim : 1));
    for (size_t n = 0; n < ndim; n++) dbf->index_lists[n] = (size_t *)std::malloc(sizeof(size_t) * (nnz ? nnz : 1));
    dbf->vals = (Ty *)std::malloc(sizeof(Ty) * (nnz ? nnz : 1));
    dbf->dims = (size_t *)std::malloc(sizeof(size_t) * (ndim ? ndim : 1));
    for (size_t n = 0; n < ndim; n++) dbf->dims[n] = 0;
    size_t k = 0, line = 0;
    for (const char *p = buf; p < end;) {
        const char *e = p;
        while (e < end && *e != '\n') ++e;
        ++line;
        diana::tns_tokens(p, e, tok);
        if (!diana::tns_skip(tok)) {
            if (tok.size() != ndim + 1)
                diana::tns_fail(filename, line, std::to_string(tok.size()) + " fields, expected " + std::to_string(ndim + 1));
            for (size_t n = 0; n < ndim; n++) {
                const size_t i = diana::tns_index(tok[n], filename, line) - 1;
                dbf->index_lists[n][k] = i;
                if (dbf->dims[n] < i + 1) dbf->dims[n] = i + 1;
            }
            dbf->vals[k] = (Ty)diana::tns_value(tok[ndim], filename, line);
            k++;
        }
        p = e < end ? e + 1 : end;
    }
    if (buf) munmap((void *)buf, size);
    close(fd);
}
#endif
