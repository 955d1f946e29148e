"""libshalom_amd -- LibShalom's small / irregular GEMM API on MI355X (gfx950).

Host-side mirror of the reference's C interface (NN_LIB/LibShalom.h:22-47)
over the extern "C" library ``libshalom.so`` built from ``csrc/``.  The names,
argument order (C first, then A, B, then M, N, K) and flags (NoTrans = 0,
Trans = 1) are the reference's; operands are device tensors (torch) or raw
device pointers (int).  There is no CPU fallback: if the HIP library is not
built, importing this module raises.

Errors: the reference functions return void and never report anything; here
the C library records a status per call and these wrappers raise
``LibShalomError`` when it is not LIBSHALOM_SUCCESS.
"""
from __future__ import annotations

import ctypes as ct
import os

__all__ = [
    "NoTrans", "Trans", "FP32", "BF16X3", "BF16", "BF16X6", "FP64",
    "LibShalomError", "lib", "sgemm", "sgemm_mp", "dgemm", "sgemm_ex", "sgemm_mp_ex",
    "dgemm_ex", "sgemm_general", "sgemm_general_noalloc", "dgemm_general", "set_thread_nums", "set_thread_ways", "get_thread_config",
    "set_precision", "get_precision", "BatchedPlan", "sgemm_batched", "dgemm_batched",
    "sgemm_backward", "sgemm_backward_strided", "sgemm_strided_batched", "autotune_batched",
    "autotune_reset", "query_tuned_tile", "query_plan", "query_batched_tile", "LIB_PATH",
]

NoTrans = 0
Trans = 1

FP32 = 0
BF16X3 = 1
BF16 = 2
BF16X6 = 3
FP64 = 4

SUCCESS = 0
STATUS_NAMES = {
    0: "LIBSHALOM_SUCCESS", 1: "LIBSHALOM_INVALID_VALUE", 2: "LIBSHALOM_INVALID_POINTER",
    3: "LIBSHALOM_HIP_ERROR", 4: "LIBSHALOM_ALLOC_FAILED", 5: "LIBSHALOM_NOT_SUPPORTED",
    6: "LIBSHALOM_NO_DEVICE",
}

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshalom.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "LibShalom.h")


class LibShalomError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)} "
                         f"({_lib.LibShalom_status_string(status).decode()})")


_L = ct.c_long
_I = ct.c_int
_P = ct.c_void_p
_LP = ct.POINTER(ct.c_long)

# (name, restype, argtypes) for every symbol declared in include/LibShalom.h.
SIGNATURES = {
    "LibShalom_sgemm": (None, [_I, _I, _P, _P, _P, _L, _L, _L]),
    "LibShalom_sgemm_mp": (None, [_I, _I, _P, _P, _P, _L, _L, _L]),
    "LibShalom_dgemm": (None, [_I, _I, _P, _P, _P, _L, _L, _L]),
    "LibShalom_set_thread_nums": (None, [_I]),
    "LibShalom_set_thread_ways": (None, [_I, _I]),
    "LibShalom_get_thread_config": (None, [ct.POINTER(_I)] * 3),
    "LibShalom_set_precision": (_I, [_I]),
    "LibShalom_get_precision": (_I, []),
    "LibShalom_get_last_error": (_I, []),
    "LibShalom_status_string": (ct.c_char_p, [_I]),
    "LibShalom_sgemm_ex": (_I, [_I, _I, _P, _P, _P, _L, _L, _L, _I, _P]),
    "LibShalom_sgemm_mp_ex": (_I, [_I, _I, _P, _P, _P, _L, _L, _L, _I, _P]),
    "LibShalom_dgemm_ex": (_I, [_I, _I, _P, _P, _P, _L, _L, _L, _P]),
    "LibShalom_sgemm_general": (_I, [_I, _I, _L, _L, _L, ct.c_float, _P, _L, _P, _L, ct.c_float, _P, _L, _I, _P]),
    "LibShalom_dgemm_general": (_I, [_I, _I, _L, _L, _L, ct.c_double, _P, _L, _P, _L, ct.c_double, _P, _L, _P]),
    "LibShalom_sgemm_general_noalloc": (_I, [_I, _I, _L, _L, _L, ct.c_float, _P, _L, _P, _L, ct.c_float, _P, _L, _I,
                                             _P]),
    "LibShalom_sgemm_batched_plan": (_I, [ct.POINTER(_P), _I, _I, _P, _P, _P, _P, _P, _P, _L, _I]),
    "LibShalom_dgemm_batched_plan": (_I, [ct.POINTER(_P), _I, _I, _P, _P, _P, _P, _P, _P, _L]),
    "LibShalom_plan_execute": (_I, [_P, _P]),
    "LibShalom_plan_destroy": (_I, [_P]),
    "LibShalom_plan_info": (_I, [_P, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double), _LP]),
    "LibShalom_sgemm_batched": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _L, _I, _P]),
    "LibShalom_dgemm_batched": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _L, _P]),
    "LibShalom_sgemm_backward": (_I, [_I, _I, _P, _P, _P, _P, _P, _L, _L, _L, _I, _P]),
    "LibShalom_sgemm_strided_batched": (_I, [_I, _I, _P, _P, _P, _L, _L, _L, _L, _L, _L, _L, _I, _P]),
    "LibShalom_sgemm_backward_strided": (_I, [_I, _I, _P, _P, _P, _P, _P, _L, _L, _L, _L, _L, _L, _L, _I, _P]),
    "LibShalom_sconv2d_forward": (_I, [_P, _P, _P] + [_L] * 13 + [_I, _P]),
    "LibShalom_conv2d_output_size": (_I, [_L] * 10 + [_LP, _LP]),
    "LibShalom_sconv2d_backward": (_I, [_P] * 5 + [_L] * 13 + [_I, _P]),
    "LibShalom_autotune_batched": (_I, [_I, _I, _I, _P, _P, _P, _L, ct.POINTER(_I), _P]),
    "LibShalom_autotune_reset": (_I, []),
    "LibShalom_query_tuned_tile": (_I, [_I, _I, _I, _L, _L, _L, ct.POINTER(_I), ct.POINTER(_I)]),
    "LibShalom_query_plan": (_I, [_I, _I, _L, _L, _L, _I] + [ct.POINTER(_I)] * 4),
    "LibShalom_query_batched_tile": (_I, [_I, _L, _L, _L, ct.POINTER(_I), ct.POINTER(_I)]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libshalom_amd: {LIB_PATH} is not built (run __graft_entry__.build() or "
            f"`make -C {_HERE}`); there is no CPU fallback")
    L = ct.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


_lib = _load()


def lib():
    return _lib


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def _ptr(x):
    """Device address of a torch tensor (must be on a GPU) or an int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if getattr(x, "device", None) is not None and x.device.type != "cuda":
            raise LibShalomError(2, "operand is not a device tensor")
        if not x.is_contiguous():
            raise ValueError("operands must be contiguous row-major tensors")
        return x.data_ptr()
    raise TypeError(f"unsupported operand type {type(x)!r}")


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _check(status: int, where: str):
    if status != SUCCESS:
        raise LibShalomError(status, where)


def _last(where: str):
    _check(_lib.LibShalom_get_last_error(), where)


# ---------------------------------------------------------------------------
# reference API (synchronous, null stream)
# ---------------------------------------------------------------------------
def _ref_ptr(x):
    """Address of an operand of the reference-API calls, which also take host
    memory as the reference does (the library stages it through the device):
    a torch tensor on the GPU or the CPU, a C-contiguous numpy array, or an int."""
    if hasattr(x, "__array_interface__") and not hasattr(x, "data_ptr"):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("operands must be contiguous row-major arrays")
        return x.ctypes.data
    if hasattr(x, "data_ptr") and getattr(x, "device", None) is not None and x.device.type == "cpu":
        if not x.is_contiguous():
            raise ValueError("operands must be contiguous row-major tensors")
        return x.data_ptr()
    return _ptr(x)


def sgemm(transa, transb, C, A, B, M, N, K):
    """LibShalom_sgemm (NN_SMM.c:5): C = op(A) op(B), C overwritten."""
    _lib.LibShalom_sgemm(transa, transb, _ref_ptr(C), _ref_ptr(A), _ref_ptr(B), M, N, K)
    _last("LibShalom_sgemm")


def sgemm_mp(transa, transb, C, A, B, M, N, K):
    """LibShalom_sgemm_mp (SMM_thread.c:75): irregular-shape path."""
    _lib.LibShalom_sgemm_mp(transa, transb, _ref_ptr(C), _ref_ptr(A), _ref_ptr(B), M, N, K)
    _last("LibShalom_sgemm_mp")


def dgemm(transa, transb, C, A, B, M, N, K):
    """LibShalom_dgemm (NN_SMM.c:30): FP64."""
    _lib.LibShalom_dgemm(transa, transb, _ref_ptr(C), _ref_ptr(A), _ref_ptr(B), M, N, K)
    _last("LibShalom_dgemm")


def set_thread_nums(num):
    """LibShalom_set_thread_nums (SMM_thread.c:116): recorded, no effect."""
    _lib.LibShalom_set_thread_nums(num)


def set_thread_ways(tm, tn):
    """LibShalom_set_thread_ways (SMM_thread.c:132): recorded, no effect."""
    _lib.LibShalom_set_thread_ways(tm, tn)


def get_thread_config():
    a, b, c = _I(), _I(), _I()
    _lib.LibShalom_get_thread_config(ct.byref(a), ct.byref(b), ct.byref(c))
    return a.value, b.value, c.value


def set_precision(p):
    _check(_lib.LibShalom_set_precision(p), "LibShalom_set_precision")


def get_precision():
    return _lib.LibShalom_get_precision()


# ---------------------------------------------------------------------------
# extended API (asynchronous on a stream)
# ---------------------------------------------------------------------------
def sgemm_ex(transa, transb, C, A, B, M, N, K, precision=FP32, stream=None):
    _check(_lib.LibShalom_sgemm_ex(transa, transb, _ptr(C), _ptr(A), _ptr(B), M, N, K, precision,
                                   _stream(stream)), "LibShalom_sgemm_ex")


def sgemm_mp_ex(transa, transb, C, A, B, M, N, K, precision=FP32, stream=None):
    _check(_lib.LibShalom_sgemm_mp_ex(transa, transb, _ptr(C), _ptr(A), _ptr(B), M, N, K, precision,
                                      _stream(stream)), "LibShalom_sgemm_mp_ex")


def dgemm_ex(transa, transb, C, A, B, M, N, K, stream=None):
    _check(_lib.LibShalom_dgemm_ex(transa, transb, _ptr(C), _ptr(A), _ptr(B), M, N, K,
                                   _stream(stream)), "LibShalom_dgemm_ex")


def sgemm_general(transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, precision=FP32, stream=None):
    """C = alpha * op(A) op(B) + beta * C, row-major, explicit leading dimensions
    (BLAS-style; no reference counterpart).  C is not read when beta is 0."""
    _check(_lib.LibShalom_sgemm_general(transa, transb, M, N, K, alpha, _ptr(A), lda, _ptr(B), ldb, beta,
                                        _ptr(C), ldc, precision, _stream(stream)), "LibShalom_sgemm_general")


def sgemm_general_noalloc(transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, precision=FP32,
                          stream=None):
    """sgemm_general without any device allocation: its split-K scratch is a
    static device array per GPU (calls that overlap on one GPU race)."""
    _check(_lib.LibShalom_sgemm_general_noalloc(transa, transb, M, N, K, alpha, _ptr(A), lda, _ptr(B), ldb, beta,
                                                _ptr(C), ldc, precision, _stream(stream)),
           "LibShalom_sgemm_general_noalloc")


def dgemm_general(transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, stream=None):
    """FP64 counterpart of sgemm_general (FP64 MFMA)."""
    _check(_lib.LibShalom_dgemm_general(transa, transb, M, N, K, alpha, _ptr(A), lda, _ptr(B), ldb, beta,
                                        _ptr(C), ldc, _stream(stream)), "LibShalom_dgemm_general")


def sgemm_backward(transa, transb, dC, A, B, dA, dB, M, N, K, precision=FP32, stream=None):
    """dA = dC op(B)^T and dB = op(A)^T dC (each stored like its operand), one launch."""
    _check(_lib.LibShalom_sgemm_backward(transa, transb, _ptr(dC), _ptr(A), _ptr(B), _ptr(dA), _ptr(dB),
                                         M, N, K, precision, _stream(stream)), "LibShalom_sgemm_backward")


def sgemm_strided_batched(transa, transb, C, A, B, M, N, K, batch, strideA, strideB, strideC, precision=FP32,
                          stream=None):
    """C_i = op(A_i) op(B_i) for a strided batch, one launch; stride 0 shares an operand (or, for C,
    sums over the batch)."""
    _check(_lib.LibShalom_sgemm_strided_batched(transa, transb, _ptr(C), _ptr(A), _ptr(B), M, N, K, batch,
                                                strideA, strideB, strideC, precision, _stream(stream)),
           "LibShalom_sgemm_strided_batched")


def sgemm_backward_strided(transa, transb, dC, A, B, dA, dB, M, N, K, batch, strideA, strideB, strideC,
                           precision=FP32, stream=None):
    """Strided-batch backward; a zero stride marks an operand shared by the batch (its gradient
    is summed over the images)."""
    _check(_lib.LibShalom_sgemm_backward_strided(transa, transb, _ptr(dC), _ptr(A), _ptr(B), _ptr(dA),
                                                 _ptr(dB), M, N, K, batch, strideA, strideB, strideC,
                                                 precision, _stream(stream)),
           "LibShalom_sgemm_backward_strided")


def conv2d_output_size(H, W, R, S, stride=(1, 1), padding=(0, 0), dilation=(1, 1)):
    """(P, Q) of a convolution of an H x W input with an R x S filter."""
    P, Q = ct.c_long(), ct.c_long()
    _check(_lib.LibShalom_conv2d_output_size(H, W, R, S, stride[0], stride[1], padding[0], padding[1],
                                             dilation[0], dilation[1], ct.byref(P), ct.byref(Q)),
           "LibShalom_conv2d_output_size")
    return P.value, Q.value


def sconv2d_forward(Y, Wt, X, batch, C, H, W, Cout, R, S, stride=(1, 1), padding=(0, 0), dilation=(1, 1),
                    precision=FP32, stream=None):
    """Y[b] = conv2d(X[b], Wt) (NCHW input, OIHW filter) as an implicit GEMM, one launch; the
    im2col matrix is gathered by the operand loads, never stored."""
    _check(_lib.LibShalom_sconv2d_forward(_ptr(Y), _ptr(Wt), _ptr(X), batch, C, H, W, Cout, R, S, stride[0],
                                          stride[1], padding[0], padding[1], dilation[0], dilation[1],
                                          precision, _stream(stream)), "LibShalom_sconv2d_forward")


def sconv2d_backward(dX, dW, dY, Wt, X, batch, C, H, W, Cout, R, S, stride=(1, 1), padding=(0, 0),
                     dilation=(1, 1), precision=FP32, stream=None):
    """dX and dW of sconv2d_forward in one launch (either may be None to skip it; dX takes
    strides up to 8, fractionally strided beyond 1); the im2col matrix is gathered, never
    stored."""
    _check(_lib.LibShalom_sconv2d_backward(_ptr(dX), _ptr(dW), _ptr(dY), _ptr(Wt), _ptr(X), batch, C, H, W,
                                           Cout, R, S, stride[0], stride[1], padding[0], padding[1],
                                           dilation[0], dilation[1], precision, _stream(stream)),
           "LibShalom_sconv2d_backward")


def _arrays(Cs, As, Bs, Ms, Ns, Ks):
    n = len(Cs)
    if not (len(As) == len(Bs) == len(Ms) == len(Ns) == len(Ks) == n):
        raise ValueError("batched arrays must have equal length")
    P = ct.c_void_p * max(n, 1)
    Lr = ct.c_long * max(n, 1)
    return (n, P(*[_ptr(c) for c in Cs]), P(*[_ptr(a) for a in As]), P(*[_ptr(b) for b in Bs]),
            Lr(*[int(m) for m in Ms]), Lr(*[int(x) for x in Ns]), Lr(*[int(k) for k in Ks]))


class BatchedPlan:
    """A batched-GEMM plan: size-bucketed work descriptors on the device and
    the launches captured in a hipGraph.  ``execute`` replays it."""

    def __init__(self, transa, transb, Cs, As, Bs, Ms, Ns, Ks, precision=FP32, fp64=False):
        n, cp, ap, bp, mp, np_, kp = _arrays(Cs, As, Bs, Ms, Ns, Ks)
        self._keep = (cp, ap, bp, mp, np_, kp)
        h = ct.c_void_p()
        if fp64:
            st = _lib.LibShalom_dgemm_batched_plan(ct.byref(h), transa, transb, cp, ap, bp, mp, np_, kp, n)
        else:
            st = _lib.LibShalom_sgemm_batched_plan(ct.byref(h), transa, transb, cp, ap, bp, mp, np_, kp, n,
                                                   precision)
        _check(st, "LibShalom_*gemm_batched_plan")
        self._h = h

    def execute(self, stream=None):
        _check(_lib.LibShalom_plan_execute(self._h, _stream(stream)), "LibShalom_plan_execute")

    def info(self):
        f, b, w = ct.c_double(), ct.c_double(), ct.c_long()
        _check(_lib.LibShalom_plan_info(self._h, ct.byref(f), ct.byref(b), ct.byref(w)), "LibShalom_plan_info")
        return {"flops": f.value, "bytes": b.value, "work_items": w.value}

    def close(self):
        if getattr(self, "_h", None):
            _lib.LibShalom_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def sgemm_batched(transa, transb, Cs, As, Bs, Ms, Ns, Ks, precision=FP32, stream=None):
    n, cp, ap, bp, mp, np_, kp = _arrays(Cs, As, Bs, Ms, Ns, Ks)
    _check(_lib.LibShalom_sgemm_batched(transa, transb, cp, ap, bp, mp, np_, kp, n, precision, _stream(stream)),
           "LibShalom_sgemm_batched")


def dgemm_batched(transa, transb, Cs, As, Bs, Ms, Ns, Ks, stream=None):
    n, cp, ap, bp, mp, np_, kp = _arrays(Cs, As, Bs, Ms, Ns, Ks)
    _check(_lib.LibShalom_dgemm_batched(transa, transb, cp, ap, bp, mp, np_, kp, n, _stream(stream)),
           "LibShalom_dgemm_batched")


def autotune_batched(precision, transa, transb, Ms, Ns, Ks, stream=None):
    """Time the candidate tiles per shape class present in (Ms, Ns, Ks) and record the
    fastest for later batched plans; returns the number of classes tuned."""
    import numpy as np
    m = np.ascontiguousarray(Ms, dtype=np.int64)
    n = np.ascontiguousarray(Ns, dtype=np.int64)
    k = np.ascontiguousarray(Ks, dtype=np.int64)
    if not (len(m) == len(n) == len(k)):
        raise ValueError("Ms, Ns, Ks differ in length")
    tuned = _I(0)
    _check(_lib.LibShalom_autotune_batched(precision, transa, transb, m.ctypes.data, n.ctypes.data,
                                           k.ctypes.data, len(m), ct.byref(tuned), _stream(stream)),
           "LibShalom_autotune_batched")
    return tuned.value


def autotune_reset():
    _check(_lib.LibShalom_autotune_reset(), "LibShalom_autotune_reset")


def query_tuned_tile(precision, transa, transb, M, N, K):
    tm, tn = _I(0), _I(0)
    _check(_lib.LibShalom_query_tuned_tile(precision, transa, transb, M, N, K, ct.byref(tm), ct.byref(tn)),
           "LibShalom_query_tuned_tile")
    return tm.value, tn.value


def query_plan(precision, M, N, K, num_cu=256, mp=False):
    tm, tn, ks, kp = _I(), _I(), _I(), _I()
    _check(_lib.LibShalom_query_plan(precision, int(mp), M, N, K, num_cu, ct.byref(tm), ct.byref(tn),
                                     ct.byref(ks), ct.byref(kp)), "LibShalom_query_plan")
    return {"tile": (tm.value, tn.value), "ksplit": ks.value, "kper": kp.value}


def query_batched_tile(precision, M, N, K):
    tm, tn = _I(), _I()
    _check(_lib.LibShalom_query_batched_tile(precision, M, N, K, ct.byref(tm), ct.byref(tn)),
           "LibShalom_query_batched_tile")
    return (tm.value, tn.value)
