"""Shapes of the batched workload: many tall-skinny matrices of uneven size.

batch_shapes(count=512) is the batched configuration: 512 matrices drawn
from std::mt19937_64(seed) in a fixed order, per matrix
    u = (x >> 11) 2^-53 from one draw x,
    m = floor(exp(log 4096 + u (log 262144 - log 4096)))   (log-uniform),
    n = (8, 16, 32, 64)[y >> 62] from the next draw y,
    m rounded down to a multiple of n (the reference's n | m,
    TallShinnyQR.h:92-98).
The engine is restated here from its published definition (Matsumoto and
Nishimura; the C++ standard's seeding and tempering), and the list it gives
for seed 0 is pinned by tests/golden/normal_stream.json, written by
libstdc++ (tests/golden/gen_normal_golden.cpp).
"""
from __future__ import annotations

import math

_MASK64 = (1 << 64) - 1


class MT19937_64:
    """std::mt19937_64: degree 312, middle word 156, twist constant
    0xB5026F5AA96619E9."""

    def __init__(self, seed: int = 0):
        s = [seed & _MASK64]
        for k in range(1, 312):
            s.append((6364136223846793005 * (s[-1] ^ (s[-1] >> 62)) + k) & _MASK64)
        self.s, self.i = s, 312

    def __call__(self) -> int:
        if self.i >= 312:
            s = self.s
            for k in range(312):
                y = (s[k] & 0xFFFFFFFF80000000) | (s[(k + 1) % 312] & 0x7FFFFFFF)
                s[k] = s[(k + 156) % 312] ^ (y >> 1) ^ (0xB5026F5AA96619E9 if y & 1 else 0)
            self.i = 0
        x = self.s[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & _MASK64


def batch_shapes(count: int = 512, m_lo: int = 4096, m_hi: int = 262144, widths=(8, 16, 32, 64), seed: int = 0):
    """[(m, n)] of the batched workload (see the module docstring)."""
    if len(widths) != 4:
        raise ValueError("widths: four values, one per top-two-bit pattern of a draw")
    eng = MT19937_64(seed)
    lo, hi = math.log(m_lo), math.log(m_hi)
    out = []
    for _ in range(count):
        u = (eng() >> 11) * 2.0 ** -53
        m = int(math.floor(math.exp(lo + u * (hi - lo))))
        n = int(widths[eng() >> 62])
        out.append((m - m % n, n))
    return out
