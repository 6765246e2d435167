"""Restatement of glibc rand() (TYPE_3 additive feedback generator) and of the
reference's polar-method normal, Util::randn (src/util.cpp:10-18).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

glibc's srand/rand (random_r.c, degree 31, separation 3): state r[0] = seed,
r[i] = 16807 r[i-1] mod (2^31 - 1) for i = 1..30 (Schrage's method),
r[i] = r[i-31] for i = 31..33, r[i] = r[i-31] + r[i-3] (mod 2^32) afterwards,
and the k-th output is r[k + 344] >> 1.  Pinned against the C library's own
rand() by tests/test_oracle_cpu.py; util_randn is pinned bit for bit against the
reference's own Util::randn, compiled from src/util.cpp into oracle/_ref/ref_randn
by oracle/Makefile (10^4 draws per seed) and recorded in tests/golden/ref_randn.json.
"""
import math

RAND_MAX = 2147483647


class GlibcRand:
    def __init__(self, seed: int):
        seed &= 0xFFFFFFFF
        if seed == 0:
            seed = 1
        r = [0] * 34
        # glibc stores the seed as int32_t
        r[0] = seed - (1 << 32) if seed >= (1 << 31) else seed
        for i in range(1, 31):
            prev = r[i - 1]
            hi = int(prev / 127773)          # C division truncates toward zero
            lo = prev - hi * 127773
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            r[i] = word
        for i in range(31, 34):
            r[i] = r[i - 31]
        self.r = [x & 0xFFFFFFFF for x in r]
        # discard outputs up to index 343
        for _ in range(34, 344):
            self._next_raw()

    def _next_raw(self) -> int:
        r = self.r
        v = (r[-31] + r[-3]) & 0xFFFFFFFF
        r.append(v)
        if len(r) > 64:
            del r[:len(r) - 40]
        return v

    def rand(self) -> int:
        return self._next_raw() >> 1


def util_randn(g: GlibcRand) -> float:
    """Util::randn (src/util.cpp:10-18): polar method, returns u*c, discards v."""
    while True:
        u = (g.rand() / RAND_MAX) * 2 - 1
        v = (g.rand() / RAND_MAX) * 2 - 1
        r = u * u + v * v
        if r == 0 or r > 1:
            continue
        c = math.sqrt(-2 * math.log(r) / r)
        return u * c


def randn_fill(g: GlibcRand, n: int):
    """Operator<double>::randn (operator_double.cpp:66-72) with one thread:
    A[i] = Util::randn() for i = 0..n-1 in order."""
    return [util_randn(g) for _ in range(n)]
