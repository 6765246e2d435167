// Util (include/util.hpp, src/util.cpp:10-30).  Util::randn is the polar
// method on the C library's rand(), returning u*c and discarding v, exactly as
// src/util.cpp:10-18 does, so srand(seed) reproduces the reference's starts.
#ifndef DIANA_AMD_UTIL_HPP
#define DIANA_AMD_UTIL_HPP
#include <cmath>
#include <cstdlib>
#include <cstring>
#include "def.hpp"

namespace Util {
inline void memcpy(void *dst, void *src, size_t len) { std::memcpy(dst, src, len); }
inline double randn() {
    for (;;) {
        double u = ((double)rand() / (RAND_MAX)) * 2 - 1;
        double v = ((double)rand() / (RAND_MAX)) * 2 - 1;
        double r = u * u + v * v;
        if (r == 0 || r > 1) continue;
        double c = std::sqrt(-2 * std::log(r) / r);
        return u * c;
    }
}
inline size_t calc_size(const shape_t &shape) {
    size_t r = 1;
    for (auto d : shape) r *= d;
    return r;
}
}  // namespace Util
#endif
