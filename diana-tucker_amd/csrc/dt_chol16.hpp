// Order-16 Cholesky with R^{-1} on the 64 lanes of one wave, and the lane
// shuffles of doubles it is built from; shared by the dense ALS update and
// the sparse factor update.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

// v_readlane of a double: lane `lane`'s value, uniform across the wave
__device__ __forceinline__ double rl_f64(double v, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// ---- order-16 Cholesky with R^{-1} on all 64 lanes of one wave ---------------
// Lane 4k + q holds rows q, q+4, q+8, q+12 of column k of the working matrix
// and of the identity under the same row operations.  Pivot j subtracts
// (G'[j][i] / G'[j][j]) times row j from every row i > j and scales row j by
// G'[j][j]^{-1/2}: the working matrix becomes R, and the product E of these
// operations satisfies E G = R, i.e. E = R^{-T}, so the identity's column k
// ends as row k of R^{-1} (no back substitution).  Row j reaches lane (k, q)
// by symmetry of the trailing block: G'[j][4s+q] = G'[4s+q][j] is slot s of
// lane 4j + q (ds_bpermute); G'[j][k] and the identity's entry come from
// lane 4k + (j mod 4) of the lane's own quad (DPP quad_perm); the pivot by
// readlane.  G'[j][j]^{-1/2}: the hardware reciprocal square root refined by
// two Newton steps (pivots below the smallest normal double take the IEEE
// square root and division).
__device__ __forceinline__ double quad_bcast_f64(double v, int src) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = (int)(uint32_t)u, hi = (int)(uint32_t)(u >> 32);
    int rlo, rhi;
    switch (src) {  // constant after unrolling: quad_perm [src, src, src, src]
        case 0: rlo = __builtin_amdgcn_mov_dpp(lo, 0x00, 0xf, 0xf, false); rhi = __builtin_amdgcn_mov_dpp(hi, 0x00, 0xf, 0xf, false); break;
        case 1: rlo = __builtin_amdgcn_mov_dpp(lo, 0x55, 0xf, 0xf, false); rhi = __builtin_amdgcn_mov_dpp(hi, 0x55, 0xf, 0xf, false); break;
        case 2: rlo = __builtin_amdgcn_mov_dpp(lo, 0xaa, 0xf, 0xf, false); rhi = __builtin_amdgcn_mov_dpp(hi, 0xaa, 0xf, 0xf, false); break;
        default: rlo = __builtin_amdgcn_mov_dpp(lo, 0xff, 0xf, 0xf, false); rhi = __builtin_amdgcn_mov_dpp(hi, 0xff, 0xf, 0xf, false); break;
    }
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)rhi << 32) | (uint32_t)rlo);
}

__device__ __forceinline__ double bpermute_f64(double v, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_ds_bpermute(lane * 4, (int)(uint32_t)u);
    const int hi = __builtin_amdgcn_ds_bpermute(lane * 4, (int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// G (R x R, LDS, R <= 16) -> Ri = chol(G + shift)^{-1} (16 x 16,
// identity-padded, [row][column]); wave 0, all of its lanes active.
__device__ __forceinline__ void chol16_rinv(double (*G)[17], double (*Ri)[17], int R, int I, int shifted,
                                            int *flag) {
    const int tid = threadIdx.x;
    if (tid >= 64) return;
    double shift = 0.0;
    if (shifted) {
        double tr = 0.0;
        for (int i = 0; i < R; i++) tr += G[i][i];
        shift = 11.0 * ((double)I * R + (double)R * (R + 1)) * 1.1102230246251565e-16 * tr;
    }
    const int k = tid >> 2, q = tid & 3;
    double col[4], z[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int i = 4 * s + q;
        col[s] = (i < R && k < R) ? G[i][k] + (i == k ? shift : 0.0) : (i == k ? 1.0 : 0.0);
        z[s] = i == k ? 1.0 : 0.0;
    }
    int bad = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int sj = j >> 2, qj = j & 3;
        double r[4];
#pragma unroll
        for (int s = 0; s < 4; s++) r[s] = (4 * s + 3 > j) ? bpermute_f64(col[s], 4 * j + q) : 0.0;
        double d = rl_f64(col[sj], 4 * j + qj);
        if (!(d > 0.0) || !isfinite(d)) {
            if (!bad) bad = j + 1;
            d = 1.0;
        }
        const double gjk = quad_bcast_f64(col[sj], qj), zjk = quad_bcast_f64(z[sj], qj);
        double ir, rjj;
        if (d >= 2.2250738585072014e-308) {
            double y = __builtin_amdgcn_rsq(d);
            double e = __builtin_fma(-d * y, y, 1.0);
            y = __builtin_fma(0.5 * y, e, y);
            e = __builtin_fma(-d * y, y, 1.0);
            y = __builtin_fma(0.5 * y, e, y);
            ir = y;
            rjj = d * y;
        } else {
            rjj = sqrt(d);
            ir = 1.0 / rjj;
        }
        const double ir2 = ir * ir, f = gjk * ir2, t = zjk * ir2;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            if (4 * s + 3 <= j) continue;
            const bool below = 4 * s + q > j;
            const double cn = col[s] - r[s] * f, zn = z[s] - r[s] * t;
            col[s] = below ? cn : col[s];
            z[s] = below ? zn : z[s];
        }
        if (q == qj) {
            col[sj] = (k == j) ? rjj : gjk * ir;
            z[sj] = zjk * ir;
        }
    }
    if (bad && tid == 0) atomicOr(flag, 4);
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int i = 4 * s + q;
        Ri[k][i] = (i < R && k < R) ? z[s] : (i == k ? 1.0 : 0.0);
    }
}
