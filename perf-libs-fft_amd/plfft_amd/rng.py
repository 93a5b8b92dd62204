"""Seeded, counter-based synthetic inputs (SplitMix64), bit-identical on any host.

Stream i of seed s is splitmix64 with state s + (i+1)*0x9e3779b97f4a7c15; the
scalars of a batch are drawn in batch-major, element-major order, real part
before imaginary part.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform_f32(seed: int, count: int) -> np.ndarray:
    """uniform[-1, 1): 2*(u24 * 2^-24) - 1, exact in float32."""
    u = (splitmix64(seed, count) >> np.uint64(40)).astype(np.float32)
    return (np.float32(2.0) * (u * np.float32(2.0 ** -24)) - np.float32(1.0)).astype(np.float32)


def uniform_f64(seed: int, count: int) -> np.ndarray:
    """uniform[-1, 1): 2*(u53 * 2^-53) - 1, exact in float64."""
    u = (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64)
    return 2.0 * (u * 2.0 ** -53) - 1.0


def normal_f64(seed: int, count: int) -> np.ndarray:
    """N(0,1) by Box-Muller in float64 (two draws per output pair)."""
    pairs = (count + 1) // 2
    z = splitmix64(seed, 2 * pairs)
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (z[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * pairs, np.float64)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:count]


def complex_uniform(seed: int, count: int, dtype=np.complex64) -> np.ndarray:
    if dtype == np.complex64:
        return uniform_f32(seed, 2 * count).view(np.complex64)
    return uniform_f64(seed, 2 * count).view(np.complex128)


def complex_normal(seed: int, count: int, dtype=np.complex128) -> np.ndarray:
    return normal_f64(seed, 2 * count).view(np.complex128).astype(dtype)


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    f = 3
    while f * f <= n:
        if n % f == 0:
            return False
        f += 2
    return True


def mixed_lengths(seed: int, count: int, lo: int = 16, hi: int = 65536,
                  prime_fraction: float = 0.1) -> list[int]:
    """Lengths log-uniform in [lo, hi]; a `prime_fraction` of them (chosen by
    a second uniform draw) are moved to the next prime, or to the largest
    prime <= hi when the next one exceeds hi."""
    u = (splitmix64(seed, 2 * count) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    out = []
    for i in range(count):
        v = int(round(lo * (hi / lo) ** u[2 * i]))
        v = min(max(v, lo), hi)
        if u[2 * i + 1] < prime_fraction:
            p = v
            while not is_prime(p):
                p += 1
            if p > hi:
                p = hi
                while not is_prime(p):
                    p -= 1
            v = p
        out.append(v)
    return out
