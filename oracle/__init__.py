"""CPU oracle for LibShalom's small / irregular GEMM path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline).  The shipped library (``libshalom_amd``)
never imports it, and the GPU path fails loudly instead of falling back here.

The C restatement lives in ``oracle.c`` / ``cpu_baseline.c`` (built into
``liboracle.so`` by ``oracle/Makefile``); see ``oracle.h`` for the reference
file:line each function follows and how it is pinned.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

U_F32 = 2.0 ** -24
U_F64 = 2.0 ** -53

_fp = ct.POINTER(ct.c_float)
_dp = ct.POINTER(ct.c_double)
_u64p = ct.POINTER(ct.c_uint64)
_L = ct.c_long


def build() -> str:
    """Compile liboracle.so with make (gcc only)."""
    subprocess.run(["make", "-B", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        L.oracle_seed_default.argtypes = [_u64p]
        L.oracle_drand48.argtypes = [_u64p]
        L.oracle_drand48.restype = ct.c_double
        L.oracle_drand48_fill.argtypes = [_u64p, _dp, _L]
        L.oracle_random_matrix.argtypes = [_u64p, _L, _L, _fp]
        L.oracle_random_matrix_d.argtypes = [_u64p, _L, _L, _dp]
        L.oracle_hello_fill.argtypes = [_u64p, _L, _L, _L, _fp, _fp]
        L.oracle_sgemm_ref.argtypes = [ct.c_int, ct.c_int, _fp, _fp, _dp, _dp, _L, _L, _L]
        L.oracle_dgemm_ref.argtypes = [ct.c_int, ct.c_int, _dp, _dp, _dp, _dp, _L, _L, _L]
        L.oracle_check_f32.argtypes = [_fp, _dp, _dp, _L, _L, ct.c_double, _dp]
        L.oracle_check_f32.restype = _L
        L.oracle_check_f64.argtypes = [_dp, _dp, _dp, _L, _L, ct.c_double, _dp]
        L.oracle_check_f64.restype = _L
        L.oracle_sgemm_ref_flat.argtypes = [ct.c_int, ct.c_int, _L] + [ct.c_void_p] * 11
        L.oracle_dgemm_ref_flat.argtypes = [ct.c_int, ct.c_int, _L] + [ct.c_void_p] * 11
        L.oracle_check_kv.argtypes = [ct.c_void_p, ct.c_int, _dp, _dp, _dp, _L, ct.c_double,
                                      ct.c_double, _dp]
        L.oracle_check_kv.restype = _L
        L.oracle_sconv2d_ref.argtypes = [_fp, _fp, _dp, _dp] + [_L] * 15
        L.oracle_sconv2d_bwd_ref.argtypes = [_fp, _fp, _fp, _dp, _dp, _dp, _dp] + [_L] * 15
        L.cpu_sgemm.argtypes = [ct.c_int, ct.c_int, _fp, _fp, _fp, _L, _L, _L]
        L.cpu_dgemm.argtypes = [ct.c_int, ct.c_int, _dp, _dp, _dp, _L, _L, _L]
        L.cpu_sgemm_batched.argtypes = [ct.c_int, ct.c_int, _L, ct.c_void_p, ct.c_void_p,
                                        ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.cpu_dgemm_batched.argtypes = L.cpu_sgemm_batched.argtypes
        L.cpu_num_threads.restype = ct.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(t)


class Drand48:
    """drand48 stream; by default from glibc's unseeded state (X = 0), which is
    what the reference harnesses use on Linux (they never call srand48)."""

    def __init__(self, state: int | None = None):
        self.state = np.zeros(1, dtype=np.uint64)
        if state is None:
            lib().oracle_seed_default(_p(self.state, _u64p))
        else:
            self.state[0] = state

    def doubles(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.float64)
        lib().oracle_drand48_fill(_p(self.state, _u64p), _p(out, _dp), n)
        return out

    def matrix(self, m: int, n: int) -> np.ndarray:
        """``random_matrix(m, n, a)``: 2*(float)drand48()-1, row-major, float32."""
        out = np.empty((m, n), dtype=np.float32)
        lib().oracle_random_matrix(_p(self.state, _u64p), m, n, _p(out, _fp))
        return out

    def matrix_d(self, m: int, n: int) -> np.ndarray:
        out = np.empty((m, n), dtype=np.float64)
        lib().oracle_random_matrix_d(_p(self.state, _u64p), m, n, _p(out, _dp))
        return out


def hello_inputs(M: int = 80, N: int = 80, K: int = 80):
    """Hello.c's A (M x K) and B buffer (K*N floats, filled as B[i*K+j])."""
    r = Drand48()
    A = np.empty(M * K, dtype=np.float32)
    B = np.empty(K * N, dtype=np.float32)
    lib().oracle_hello_fill(_p(r.state, _u64p), M, N, K, _p(A, _fp), _p(B, _fp))
    return A, B


def a_shape(transa: int, M: int, K: int):
    return (K, M) if transa else (M, K)


def b_shape(transb: int, N: int, K: int):
    return (N, K) if transb else (K, N)


def sgemm_ref(transa, transb, A, B, M, N, K):
    """Long-double reference of C = op(A) op(B) for float32 inputs.
    Returns (Cref float64 [M,N], absAB float64 [M,N])."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    assert A.size == M * K and B.size == K * N
    C = np.empty((M, N), dtype=np.float64)
    S = np.empty((M, N), dtype=np.float64)
    lib().oracle_sgemm_ref(transa, transb, _p(A, _fp), _p(B, _fp), _p(C, _dp), _p(S, _dp), M, N, K)
    return C, S


def dgemm_ref(transa, transb, A, B, M, N, K):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    assert A.size == M * K and B.size == K * N
    C = np.empty((M, N), dtype=np.float64)
    S = np.empty((M, N), dtype=np.float64)
    lib().oracle_dgemm_ref(transa, transb, _p(A, _dp), _p(B, _dp), _p(C, _dp), _p(S, _dp), M, N, K)
    return C, S


def conv_out(H, R, stride, pad, dil):
    """Output size of a convolution along one axis (0 when the dilated filter does not fit)."""
    span = H + 2 * pad - dil * (R - 1) - 1
    return 0 if span < 0 else span // stride + 1


def sconv2d_ref(X, Wt, stride=(1, 1), padding=(0, 0), dilation=(1, 1)):
    """Long-double reference of the convolution of X (batch, C, H, W) with Wt (Cout, C, R, S).
    Returns (Yref float64 [batch, Cout, P, Q], absWX float64, same shape)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    Wt = np.ascontiguousarray(Wt, dtype=np.float32)
    batch, C, H, W = X.shape
    Cout, C2, R, S = Wt.shape
    assert C2 == C
    P = conv_out(H, R, stride[0], padding[0], dilation[0])
    Q = conv_out(W, S, stride[1], padding[1], dilation[1])
    Y = np.empty((batch, Cout, P, Q), dtype=np.float64)
    A = np.empty_like(Y)
    lib().oracle_sconv2d_ref(_p(X, _fp), _p(Wt, _fp), _p(Y, _dp), _p(A, _dp), batch, C, H, W, Cout, R, S,
                             stride[0], stride[1], padding[0], padding[1], dilation[0], dilation[1], P, Q)
    return Y, A


def sconv2d_bwd_ref(X, Wt, dY, stride=(1, 1), padding=(0, 0), dilation=(1, 1)):
    """Long-double gradients of the convolution for the loss sum(Y * dY).
    Returns (dX, |dX| bound sums, dW, |dW| bound sums), float64, shaped like X and Wt."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    Wt = np.ascontiguousarray(Wt, dtype=np.float32)
    dY = np.ascontiguousarray(dY, dtype=np.float32)
    batch, C, H, W = X.shape
    Cout, _, R, S = Wt.shape
    P = conv_out(H, R, stride[0], padding[0], dilation[0])
    Q = conv_out(W, S, stride[1], padding[1], dilation[1])
    assert dY.shape == (batch, Cout, P, Q)
    dX = np.empty(X.shape, dtype=np.float64)
    aX = np.empty_like(dX)
    dW = np.empty(Wt.shape, dtype=np.float64)
    aW = np.empty_like(dW)
    lib().oracle_sconv2d_bwd_ref(_p(X, _fp), _p(Wt, _fp), _p(dY, _fp), _p(dX, _dp), _p(aX, _dp), _p(dW, _dp),
                                 _p(aW, _dp), batch, C, H, W, Cout, R, S, stride[0], stride[1], padding[0],
                                 padding[1], dilation[0], dilation[1], P, Q)
    return dX, aX, dW, aW


def check(C, Cref, absAB, K, u=U_F32):
    """Elementwise |C - Cref| <= 2*K*u*absAB.  Returns (violations, stats)."""
    st = np.zeros(4, dtype=np.float64)
    Cref = np.ascontiguousarray(Cref, dtype=np.float64)
    absAB = np.ascontiguousarray(absAB, dtype=np.float64)
    if C.dtype == np.float64:
        C = np.ascontiguousarray(C)
        bad = lib().oracle_check_f64(_p(C, _dp), _p(Cref, _dp), _p(absAB, _dp), C.size, K, u, _p(st, _dp))
    else:
        C = np.ascontiguousarray(C, dtype=np.float32)
        bad = lib().oracle_check_f32(_p(C, _fp), _p(Cref, _dp), _p(absAB, _dp), C.size, K, u, _p(st, _dp))
    return int(bad), {"max_ratio": st[0], "max_rel": st[1], "mean_rel": st[2], "max_abs": st[3]}


def gemm_ref_flat(transa, transb, A, B, offA, offB, offC, Ms, Ns, Ks, total_c):
    """References for many GEMMs packed back to back in flat arrays.
    Returns (Cref, absAB, Kelem), each of length total_c."""
    f64 = A.dtype == np.float64
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B)
    arrs = [np.ascontiguousarray(x, dtype=np.int64) for x in (offA, offB, offC, Ms, Ns, Ks)]
    Cref = np.zeros(total_c, dtype=np.float64)
    S = np.zeros(total_c, dtype=np.float64)
    Ke = np.zeros(total_c, dtype=np.float64)
    fn = lib().oracle_dgemm_ref_flat if f64 else lib().oracle_sgemm_ref_flat
    fn(transa, transb, len(arrs[3]), A.ctypes.data, B.ctypes.data, *[x.ctypes.data for x in arrs],
       Cref.ctypes.data, S.ctypes.data, Ke.ctypes.data)
    return Cref, S, Ke


def check_kv(C, Cref, absAB, Kelem, u=U_F32, extra=0.0):
    """Elementwise |C - Cref| <= (2*K_i*u + extra)*absAB_i with per-element K."""
    st = np.zeros(4, dtype=np.float64)
    C = np.ascontiguousarray(C)
    is64 = int(C.dtype == np.float64)
    if not is64:
        C = np.ascontiguousarray(C, dtype=np.float32)
    bad = lib().oracle_check_kv(C.ctypes.data, is64, _p(np.ascontiguousarray(Cref), _dp),
                                _p(np.ascontiguousarray(absAB), _dp), _p(np.ascontiguousarray(Kelem), _dp),
                                C.size, u, extra, _p(st, _dp))
    return int(bad), {"max_ratio": st[0], "max_rel": st[1], "mean_rel": st[2], "max_abs": st[3]}


def cpu_sgemm(transa, transb, A, B, M, N, K):
    C = np.empty((M, N), dtype=np.float32)
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    lib().cpu_sgemm(transa, transb, _p(C, _fp), _p(A, _fp), _p(B, _fp), M, N, K)
    return C


def cpu_dgemm(transa, transb, A, B, M, N, K):
    C = np.empty((M, N), dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    lib().cpu_dgemm(transa, transb, _p(C, _dp), _p(A, _dp), _p(B, _dp), M, N, K)
    return C


def cpu_gemm_batched(transa, transb, Cs, As, Bs, Ms, Ns, Ks):
    """Run the batched CPU loop over lists of numpy arrays (all float32 or all
    float64); results are written into ``Cs`` in place."""
    count = len(Cs)
    f64 = Cs[0].dtype == np.float64
    cp = np.array([c.ctypes.data for c in Cs], dtype=np.uintp)
    ap = np.array([a.ctypes.data for a in As], dtype=np.uintp)
    bp = np.array([b.ctypes.data for b in Bs], dtype=np.uintp)
    m = np.ascontiguousarray(Ms, dtype=np.int64)
    n = np.ascontiguousarray(Ns, dtype=np.int64)
    k = np.ascontiguousarray(Ks, dtype=np.int64)
    fn = lib().cpu_dgemm_batched if f64 else lib().cpu_sgemm_batched
    fn(transa, transb, count, cp.ctypes.data, ap.ctypes.data, bp.ctypes.data,
       m.ctypes.data, n.ctypes.data, k.ctypes.data)


def cpu_threads() -> int:
    return int(lib().cpu_num_threads())
