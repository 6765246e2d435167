"""Philox4x32-10 stream restated in numpy -- TEST INFRASTRUCTURE ONLY.

Same mapping as diana-tucker_amd/csrc/dt_gen.hip: element i of stream s under
seed k uses call q = i // 2 with counter (q_lo, q_hi, s_lo, s_hi) and key
(k_lo, k_hi); a = x0 | x1 << 32, b = x2 | x3 << 32; u1 = ((a >> 11) + 1) 2^-53,
u2 = (b >> 11) 2^-53; normal = sqrt(-2 ln u1) * (cos if i even else sin)(2 pi u2).
Round constants from Salmon et al., "Parallel random numbers: as easy as
1, 2, 3" (SC'11).  The published known-answer vectors of that paper pin it in
tests/test_oracle_cpu.py.
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1, rounds=10):
    c0 = np.asarray(c0, dtype=np.uint64) & MASK
    c1 = np.asarray(c1, dtype=np.uint64) & MASK
    c2 = np.asarray(c2, dtype=np.uint64) & MASK
    c3 = np.asarray(c3, dtype=np.uint64) & MASK
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for r in range(rounds):
        if r > 0:
            k0 = (k0 + W0) & 0xFFFFFFFF
            k1 = (k1 + W1) & 0xFFFFFFFF
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ np.uint64(k0)), lo1, (hi0 ^ c3 ^ np.uint64(k1)), lo0
    return c0, c1, c2, c3


def _pairs(seed, stream, q):
    q = np.asarray(q, dtype=np.uint64)
    x0, x1, x2, x3 = philox4x32_10(q & MASK, q >> np.uint64(32), np.uint64(stream & 0xFFFFFFFF),
                                   np.uint64(stream >> 32), seed & 0xFFFFFFFF, seed >> 32)
    a = x0 | (x1 << np.uint64(32))
    b = x2 | (x3 << np.uint64(32))
    return a, b


def normal(n, seed, stream, start=0):
    i = np.arange(start, start + n, dtype=np.uint64)
    a, b = _pairs(seed, stream, i >> np.uint64(1))
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    rad = np.sqrt(-2.0 * np.log(u1))
    th = 2.0 * np.pi * u2
    return np.where((i & np.uint64(1)) == 1, rad * np.sin(th), rad * np.cos(th))


def uniform(n, seed, stream, start=0):
    i = np.arange(start, start + n, dtype=np.uint64)
    a, b = _pairs(seed, stream, i >> np.uint64(1))
    x = np.where((i & np.uint64(1)) == 1, b, a)
    return ((x >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
