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


def drand48_stream(count: int, state: int = 0):
    """`count` values of the POSIX drand48 generator, X' = (0x5DEECE66D X +
    0xB) mod 2^48, value X' / 2^48, from `state` (0: glibc's state when the
    program never seeds it)."""
    import numpy as np
    out = np.empty(count, dtype=np.float64)
    x = state
    for i in range(count):
        x = (0x5DEECE66D * x + 0xB) & ((1 << 48) - 1)
        out[i] = x / float(1 << 48)
    return out


def harness_entries(values):
    """The reference harness's entry recipe (float)(2 * (double)(float)r - 1)
    (benchmark/small_SGEMM/LibShalom_sgemm.c:33) applied to drand48 values."""
    import numpy as np
    return (2.0 * values.astype(np.float32).astype(np.float64) - 1.0).astype(np.float32)


def hello_inputs(M: int = 80, N: int = 80, K: int = 80):
    """Hello.c's inputs: A (M x K row-major) then the B buffer filled as
    B[i*K + j] for i < K, j < N (Hello.c:20-29)."""
    import numpy as np
    v = harness_entries(drand48_stream(M * K + K * N))
    A = v[:M * K].copy()
    B = np.empty(K * N, dtype=np.float32)
    for i in range(K):
        for j in range(N):
            B[i * K + j] = v[M * K + i * N + j]
    return A, B
