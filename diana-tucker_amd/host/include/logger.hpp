// Logging conventions of the reference (include/logger.hpp:24-158): output()
// prints on rank 0 to stderr, error() reports and exits with -3, assert-style
// fatal() exits with -4.  Building with DIANA_THROW_ON_ERROR turns error() into
// a std::runtime_error so test drivers can observe failures.
#ifndef DIANA_AMD_LOGGER_HPP
#define DIANA_AMD_LOGGER_HPP
#include <iostream>
#include <stdexcept>
#include <string>

int mpi_rank();

#define output(x)                                  \
    do {                                           \
        if (mpi_rank() == 0) std::cerr << (x) << std::endl; \
    } while (0)

namespace diana_detail {
[[noreturn]] inline void raise_error(const std::string &msg, const char *file, int line) {
    std::cerr << "[ERROR " << file << ":" << line << "] " << msg << std::endl;
#ifdef DIANA_THROW_ON_ERROR
    throw std::runtime_error(msg);
#else
    std::exit(-3);
#endif
}
}  // namespace diana_detail

#define error(x) diana_detail::raise_error((x), __FILE__, __LINE__)
#endif
