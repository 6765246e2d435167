// Driver for the reference's own Util::randn (src/util.cpp:10-18), compiled
// from the reference sources by oracle/Makefile into oracle/_ref/ref_randn.
// Test infrastructure only: it prints, for `srand(seed)`, the first `count`
// draws of Util::randn() as the 16 hex digits of their IEEE-754 bits, one per
// line, so tests can compare the restatements bit for bit.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "util.hpp"

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s seed count\n", argv[0]);
        return 2;
    }
    const unsigned seed = (unsigned)std::strtoul(argv[1], nullptr, 10);
    const long count = std::strtol(argv[2], nullptr, 10);
    srand(seed);
    for (long i = 0; i < count; i++) {
        const double x = Util::randn();
        uint64_t bits;
        std::memcpy(&bits, &x, sizeof bits);
        std::printf("%016llx\n", (unsigned long long)bits);
    }
    return 0;
}
