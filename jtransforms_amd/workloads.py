"""Synthetic workloads of the benchmark configurations (shared by bench.py and
the tests; no dependence on the oracle)."""
from __future__ import annotations

import math

import numpy as np


def _smooth_numbers(lo: int, hi: int):
    out = []
    p2 = 1
    while p2 <= 4 * hi:
        p3 = p2
        while p3 <= 4 * hi:
            p5 = p3
            while p5 <= 4 * hi:
                if lo // 2 <= p5:
                    out.append(p5)
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return np.array(sorted(set(out)), dtype=np.int64)


def _is_prime(n: int) -> bool:
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    r = int(math.isqrt(n))
    for d in range(3, r + 1, 2):
        if n % d == 0:
            return False
    return True


def _nearest(cands: np.ndarray, x: float) -> int:
    i = int(np.argmin(np.abs(np.log(cands) - math.log(x))))
    return int(cands[i])


def ragged_lengths(count: int = 20000, lo: int = 16, hi: int = 65536, seed: int = 20110602) -> np.ndarray:
    """Lengths log-uniform in [lo, hi]: 70% snapped to the nearest 2^a 3^b 5^c,
    25% to the nearest power of two, 5% to the nearest prime (all within
    [lo, hi])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = np.exp(rng.uniform(math.log(lo), math.log(hi), count))
    kind = rng.uniform(0, 1, count)
    smooth = _smooth_numbers(lo, hi)
    smooth = smooth[(smooth >= lo) & (smooth <= hi)]
    pow2 = np.array([1 << k for k in range(int(math.log2(lo)), int(math.log2(hi)) + 1)], dtype=np.int64)
    out = np.empty(count, dtype=np.int64)
    for i in range(count):
        x = u[i]
        if kind[i] < 0.70:
            out[i] = _nearest(smooth, x)
        elif kind[i] < 0.95:
            out[i] = _nearest(pow2, x)
        else:
            c = int(round(x))
            d = 0
            while True:
                if lo <= c - d <= hi and _is_prime(c - d):
                    out[i] = c - d
                    break
                if lo <= c + d <= hi and _is_prime(c + d):
                    out[i] = c + d
                    break
                d += 1
    return out
