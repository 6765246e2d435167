// Process-wide device context of the host layer and the mapping of C-ABI
// status codes onto the reference's error() convention.
#ifndef DIANA_AMD_DEVICE_HPP
#define DIANA_AMD_DEVICE_HPP
#include <cstdlib>
#include <string>
#include "diana_hip.h"
#include "logger.hpp"

namespace diana {
struct Device {
    // The HIP runtime's initialisation draws from the C library's rand()
    // stream.  The reference seeds that stream (srand in main.cpp:16,
    // spmain.cpp:17) and draws its start factors from it later, so the
    // context is created on a private generator state and the caller's state
    // is put back untouched.
    static dt_ctx *ctx() {
        static dt_ctx *c = [] {
            static char priv[256];
            char *caller = initstate(1u, priv, sizeof(priv));
            dt_ctx *p = nullptr;
            int st = dt_create(&p, 0);
            setstate(caller);
            if (st != DT_OK) error("dt_create failed (no HIP device?), status " + std::to_string(st));
            return p;
        }();
        return c;
    }
};

inline void check(int st, const char *what) {
    if (st != DT_OK)
        error(std::string(what) + ": " + dt_last_error(Device::ctx()) + " (status " +
              std::to_string(st) + ")");
}

// RAII device buffer
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) : bytes(b) { check(dt_malloc(Device::ctx(), &p, b), "dt_malloc"); }
    DevBuf(const void *host, size_t b) : DevBuf(b) { check(dt_h2d(Device::ctx(), p, host, b), "dt_h2d"); }
    ~DevBuf() {
        if (p) dt_free(Device::ctx(), p);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    void to_host(void *host, size_t b) const { check(dt_d2h(Device::ctx(), host, p, b), "dt_d2h"); }
};

template <typename Ty> struct DType;
template <> struct DType<double> { static constexpr int v = DT_F64; };
template <> struct DType<float> { static constexpr int v = DT_F32; };
}  // namespace diana
#endif
