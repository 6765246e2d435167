// Basic types of the reference API (include/def.hpp:7-28), kept verbatim in meaning.
#ifndef DIANA_AMD_DEF_HPP
#define DIANA_AMD_DEF_HPP
#include <complex>
#include <cstddef>
#include <cstdlib>
#include <vector>

typedef std::vector<int> vint;
typedef std::vector<size_t> shape_t;
typedef float float32;
typedef double float64;
typedef std::complex<float> complex32;
typedef std::complex<double> complex64;

#define DIANA_CEILDIV(n, k) (((n) + (k)-1) / (k))
#define DIANA_UNUSED(x) (void)x;
#endif
