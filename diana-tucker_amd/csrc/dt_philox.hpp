// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11)
// shared by the dense Tucker generator and the sparse Zipf generator.
// Element i of stream s under seed k is a pure function of (k, s, i); see
// dt_gen.hip for the mapping to normals and uniforms.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

struct u32x4s {
    uint32_t x, y, z, w;
};

__host__ __device__ inline u32x4s philox4x32_10(u32x4s ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += W0;
            k1 += W1;
        }
        const uint64_t p0 = (uint64_t)M0 * ctr.x;
        const uint64_t p1 = (uint64_t)M1 * ctr.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        ctr = u32x4s{hi1 ^ ctr.y ^ k0, lo1, hi0 ^ ctr.w ^ k1, lo0};
    }
    return ctr;
}

__device__ inline void philox_pair(uint64_t seed, uint64_t stream, uint64_t q, uint64_t &a,
                                   uint64_t &b) {
    u32x4s c{(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
    u32x4s r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    a = (uint64_t)r.x | ((uint64_t)r.y << 32);
    b = (uint64_t)r.z | ((uint64_t)r.w << 32);
}

__device__ inline double normal_at(uint64_t seed, uint64_t stream, uint64_t i) {
    uint64_t a, b;
    philox_pair(seed, stream, i >> 1, a, b);
    const double u1 = (double)((a >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(b >> 11) * 0x1.0p-53;
    const double rad = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586476925286766559 * u2;
    return (i & 1) ? rad * sin(th) : rad * cos(th);
}

__device__ inline double uniform_at(uint64_t seed, uint64_t stream, uint64_t i) {
    uint64_t a, b;
    philox_pair(seed, stream, i >> 1, a, b);
    const uint64_t x = (i & 1) ? b : a;
    return (double)((x >> 11) + 1) * 0x1.0p-53;
}

