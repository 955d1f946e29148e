"""Synthetic workloads of the benchmark configurations (shapes only).

The reference's harnesses draw matrix entries from drand48 and use fixed
shape lists (benchmark/*/, evaluate/test.c:114-116); the batched workloads
have no reference counterpart, so their shapes are drawn here from a
splitmix64 stream (seed 0), reproducible from any language.
"""
from __future__ import annotations

MASK64 = (1 << 64) - 1


def splitmix64(seed: int = 0):
    """Infinite splitmix64 stream (Steele, Lea & Flood 2014)."""
    x = seed & MASK64
    while True:
        x = (x + 0x9E3779B97F4A7C15) & MASK64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def uniform_shapes(count: int, lo: int = 1, hi: int = 128, seed: int = 0):
    """`count` (M, N, K) triples, each entry uniform on {lo..hi}."""
    g = splitmix64(seed)
    span = hi - lo + 1
    return [(lo + next(g) % span, lo + next(g) % span, lo + next(g) % span) for _ in range(count)]


CP2K_DIMS = (4, 5, 9, 13, 16, 22, 23, 26, 32)


def cp2k_shapes(count: int, seed: int = 0):
    """`count` (M, N, K) triples drawn from the small-block dims of
    quantum-chemistry codes (CP2K_kernels/LibShalom_dgemm.c:74-76 uses 5..26)."""
    g = splitmix64(seed)
    d = CP2K_DIMS
    return [(d[next(g) % len(d)], d[next(g) % len(d)], d[next(g) % len(d)]) for _ in range(count)]


SMALL_SWEEP = [(n, n, n) for n in range(8, 129, 8)]
TALL_SKINNY = [(M, N, K) for M in (4096, 16384, 65536) for N in (16, 32, 64) for K in (64, 256, 1024)]
CONV_SHAPES = [(64, 3136, 576), (128, 784, 1152), (256, 196, 2304), (512, 49, 4608)]
