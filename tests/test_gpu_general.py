"""GPU parity tests of the BLAS-style entry points LibShalom_sgemm_general /
LibShalom_dgemm_general (explicit leading dimensions, C = alpha*op(A)op(B) +
beta*C), through the C-ABI, against the oracle's long-double product.

Bar: |C - C_ref| <= |alpha| * (2*K*u + extra) * (|op(A)||op(B)|)_ij
+ 2*u*(|alpha*P| + |beta*C0|)_ij (u = 2^-24 for FP32 output, 2^-53 for FP64;
extra = the mode's per-product split term).  Operands live inside larger
padded buffers whose padding holds sentinels: nothing outside the M x N
window of C may change, and with beta = 0 a NaN-filled C must not leak into
the result (C is not read).
"""
import numpy as np
import pytest
import torch

import libshalom_amd as ls
import oracle

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
U64 = 2.0 ** -53
EXTRA = {ls.FP32: 0.0, ls.BF16X3: 3 * 2.0 ** -16, ls.BF16X6: 0.0}
SENTINEL = 1234.5


def padded(rng, rows, cols, ld, dtype):
    """A rows x ld buffer whose first `cols` columns hold U[-1,1) data and the
    rest a sentinel; returns (buffer, data view)."""
    buf = np.full((rows, ld), SENTINEL, dtype=dtype)
    buf[:, :cols] = rng.uniform(-1.0, 1.0, (rows, cols)).astype(dtype)
    return buf, buf[:, :cols]


def run_case(dev, f64, prec, ta, tb, M, N, K, alpha, beta, pads, seed):
    rng = np.random.default_rng(seed)
    dt = np.float64 if f64 else np.float32
    ar, ac = (K, M) if ta else (M, K)
    br, bc = (N, K) if tb else (K, N)
    lda, ldb, ldc = ac + pads[0], bc + pads[1], N + pads[2]
    Abuf, A = padded(rng, ar, ac, lda, dt)
    Bbuf, B = padded(rng, br, bc, ldb, dt)
    Cbuf, C0 = padded(rng, M, N, ldc, dt)
    if beta == 0:
        Cbuf[:, :N] = np.nan  # must not be read
    dA, dB, dC = (torch.from_numpy(x.copy()).to(dev) for x in (Abuf, Bbuf, Cbuf))
    if f64:
        ls.dgemm_general(ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc)
    else:
        ls.sgemm_general(ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, prec)
    torch.cuda.synchronize()
    out = dC.cpu().numpy()
    # the padding of C is untouched
    assert np.array_equal(out[:, N:], Cbuf[:, N:])
    Cw = out[:, :N].astype(np.float64)
    if f64:
        P, absAB = oracle.dgemm_ref(ta, tb, A.copy(), B.copy(), M, N, K)
    else:
        P, absAB = oracle.sgemm_ref(ta, tb, A.copy(), B.copy(), M, N, K)
    P, absAB = P.reshape(M, N), absAB.reshape(M, N)
    c0 = np.zeros((M, N)) if beta == 0 else C0.astype(np.float64)
    ref = alpha * P + beta * c0
    u = U64 if f64 else U32
    extra = 0.0 if f64 else EXTRA[prec]
    tol = abs(alpha) * (2 * max(K, 1) * u + extra) * absAB + 2 * u * (np.abs(alpha * P) + np.abs(beta * c0)) \
        + 1e-300
    err = np.abs(Cw - ref)
    assert np.isfinite(Cw).all()
    assert (err <= tol).all(), (M, N, K, ta, tb, prec, alpha, beta, float((err / tol).max()))


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3, ls.BF16X6])
def test_sgemm_general_ld_alpha_beta(dev, ta, tb, prec):
    for i, (M, N, K) in enumerate([(37, 53, 29), (130, 70, 300), (1, 129, 7)]):
        for alpha, beta in [(1.5, 0.0), (-0.75, 0.5), (1.0, 1.0)]:
            run_case(dev, False, prec, ta, tb, M, N, K, alpha, beta, (3, 5, 7), 100 * i + ta * 10 + tb)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_dgemm_general_ld_alpha_beta(dev, ta, tb):
    for i, (M, N, K) in enumerate([(23, 26, 13), (70, 130, 200)]):
        for alpha, beta in [(2.0, 0.0), (-0.5, 0.25)]:
            run_case(dev, True, None, ta, tb, M, N, K, alpha, beta, (1, 2, 3), 7 + i)


@pytest.mark.parametrize("f64", [False, True])
def test_general_split_k_applies_alpha_beta_once(dev, f64):
    # a shape the planner splits along K: the reduction kernel applies alpha/beta
    num_cu = torch.cuda.get_device_properties(0).multi_processor_count
    M, N, K = 64, 64, 8192
    plan = ls.query_plan(ls.FP64 if f64 else ls.FP32, M, N, K, num_cu)
    assert plan["ksplit"] > 1
    for alpha, beta in [(0.5, 0.0), (2.0, -1.0)]:
        run_case(dev, f64, ls.FP32, 0, 1, M, N, K, alpha, beta, (4, 4, 4), 3)


def test_general_k0_scales_c(dev):
    # K = 0: C = beta * C (alpha multiplies an empty product)
    run_case(dev, False, ls.FP32, 0, 0, 17, 9, 0, 3.0, 0.5, (2, 2, 2), 5)
    run_case(dev, False, ls.FP32, 0, 0, 17, 9, 0, 3.0, 0.0, (2, 2, 2), 5)


def test_general_matches_reference_api_at_defaults(dev):
    # alpha 1, beta 0, natural leading dimensions: identical bits to LibShalom_sgemm
    rng = np.random.default_rng(11)
    M, N, K = 80, 80, 80
    A = torch.from_numpy(rng.uniform(-1, 1, (M, K)).astype(np.float32)).to(dev)
    B = torch.from_numpy(rng.uniform(-1, 1, (N, K)).astype(np.float32)).to(dev)
    C1 = torch.full((M, N), float("nan"), device=dev)
    C2 = torch.full((M, N), float("nan"), device=dev)
    ls.sgemm(ls.NoTrans, ls.Trans, C1, A, B, M, N, K)
    ls.sgemm_general(ls.NoTrans, ls.Trans, M, N, K, 1.0, A, K, B, K, 0.0, C2, N)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
