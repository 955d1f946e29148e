"""Writes the golden fixtures in this directory.

Independent of the oracle on purpose: inputs come from glibc's own drand48
(through ctypes, reset to the unseeded state with seed48({0,0,0})), and
expected products from numpy's float64 matmul.  LibShalom ships no golden
vectors; its harnesses draw inputs as `2.0 * (float)drand48() - 1.0` from the
unseeded stream (benchmark/small_SGEMM/LibShalom_sgemm.c:26-34, Hello.c:20-29).

Run: python tests/golden/make_golden.py
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
libc = ctypes.CDLL("libc.so.6")
libc.drand48.restype = ctypes.c_double
libc.seed48.argtypes = [ctypes.POINTER(ctypes.c_ushort)]
libc.seed48.restype = ctypes.POINTER(ctypes.c_ushort)


def reset():
    libc.seed48((ctypes.c_ushort * 3)(0, 0, 0))


def recipe(n):
    """2.0 * (float)drand48() - 1.0 into a float."""
    out = np.empty(n, dtype=np.float32)
    for i in range(n):
        f = np.float32(libc.drand48())
        out[i] = np.float32(2.0 * np.float64(f) - 1.0)
    return out


def hexs(a):
    return [float(x).hex() for x in np.asarray(a, dtype=np.float64).ravel()]


def main():
    reset()
    first = [libc.drand48() for _ in range(32)]
    reset()
    floats = recipe(16)
    with open(os.path.join(HERE, "drand48_first.json"), "w") as f:
        json.dump({"source": "glibc drand48 from the unseeded state (seed48 {0,0,0})",
                   "drand48": [x.hex() for x in first],
                   "recipe_f32": hexs(floats)}, f, indent=1)

    # Hello.c: A 80x80 row-major, then B filled as B[i*K+j]; C = A * B^T.
    M = N = K = 80
    reset()
    A = recipe(M * K).reshape(M, K)
    B = recipe(K * N).reshape(N, K)  # B[i*K + j] with N == K: the NT operand B[n][k]
    C = A.astype(np.float64) @ B.astype(np.float64).T
    with open(os.path.join(HERE, "hello_80_nt.json"), "w") as f:
        json.dump({"M": M, "N": N, "K": K, "transa": 0, "transb": 1,
                   "sum": float(C.sum()).hex(), "abs_sum": float(np.abs(C).sum()).hex(),
                   "corner": hexs(C[:4, :4]), "diag": hexs(np.diag(C))}, f, indent=1)

    # Tiny shapes, every trans combination; inputs from one continuing stream.
    reset()
    cases = []
    for (M, N, K) in [(1, 1, 1), (3, 5, 7), (5, 3, 2), (7, 1, 9), (4, 6, 1), (2, 9, 5)]:
        for ta in (0, 1):
            for tb in (0, 1):
                A = recipe(M * K)
                B = recipe(K * N)
                opA = (A.reshape(K, M).T if ta else A.reshape(M, K)).astype(np.float64)
                opB = (B.reshape(N, K).T if tb else B.reshape(K, N)).astype(np.float64)
                C = opA @ opB
                cases.append({"M": M, "N": N, "K": K, "transa": ta, "transb": tb,
                              "A": hexs(A), "B": hexs(B), "C": hexs(C)})
    with open(os.path.join(HERE, "tiny_cases.json"), "w") as f:
        json.dump(cases, f)


if __name__ == "__main__":
    main()
