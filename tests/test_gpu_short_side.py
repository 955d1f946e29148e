"""GPU parity of the short-side kernel (a tall C with N <= 64 or a wide one
with M <= 64, K >= 512), FP32 and the split-BF16 modes: every transpose
combination, ragged long and short sides, K not a multiple of the 16- (32-)
k groups or of the waves' shares, alpha / beta with padded leading
dimensions.  Each case first checks (query_plan) that the planner routes it
to the short-side kernel; every element of C is checked against a float64
product on the device (FP32 and BF16X6 to 2 K u |A||B|, BF16X3 and BF16
with their per-product terms) and, in FP32, a sample of rows against the
oracle, C prefilled with NaN (padding must stay untouched)."""
import numpy as np
import pytest
import torch

import libshalom_amd as ls
import oracle

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
TRANS = [(0, 0), (0, 1), (1, 0), (1, 1)]
# extra relative error per product |a||b| of the one- and two-piece splits
EXTRA = {ls.FP32: 0.0, ls.BF16X6: 0.0, ls.BF16X3: 3 * 2.0 ** -16, ls.BF16: 2 * 2.0 ** -8}
# (M, N, K): tall and wide, long sides not multiples of 16, short sides of
# 1 to 64 (one, two and four 16-column blocks), K ragged against 16 and 64
# (the last two take the eight-wave kernel in NN: one round of workgroups
# carrying 4096 of K x short side)
SHAPES = [(1031, 5, 777), (2048, 17, 1024), (4093, 33, 520), (3000, 64, 512), (1024, 1, 2000),
          (7, 1500, 1000), (40, 4096, 512), (64, 2047, 512), (1, 3000, 700), (4090, 20, 2048), (4085, 7, 4096)]


def is_short_side_plan(M, N, K, prec=ls.FP32):
    p = ls.query_plan(prec, M, N, K, num_cu=torch.cuda.get_device_properties(0).multi_processor_count)
    return p["ksplit"] == 1 and 16 in p["tile"] and p["tile"] not in {(32, 32), (64, 64), (128, 128)}


def run_general(dev, ta, tb, M, N, K, alpha, beta, pad, seed, prec=ls.FP32):
    r = oracle.Drand48(seed)
    lda = (M if ta else K) + pad
    ldb = (K if tb else N) + pad
    ldc = N + pad
    A = r.matrix(*((K, lda) if ta else (M, lda)))
    B = r.matrix(*((N, ldb) if tb else (K, ldb)))
    C0 = r.matrix(M, ldc)
    dA = torch.from_numpy(A).to(dev)
    dB = torch.from_numpy(B).to(dev)
    dC = torch.from_numpy(C0.copy()).to(dev)
    if beta == 0.0:
        dC[:, :N] = float("nan")  # not read when beta is 0
    ls.sgemm_general(ta, tb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, prec)
    torch.cuda.synchronize()
    C = dC.cpu().numpy()
    assert np.array_equal(C[:, N:], C0[:, N:]), "padding written"
    opA = (A[:, :M].T if ta else A[:, :K]).astype(np.float64)
    opB = (B[:, :K].T if tb else B[:, :N]).astype(np.float64)
    # every element against a float64 product on the device
    tA = torch.from_numpy(np.ascontiguousarray(opA)).to(dev)
    tB = torch.from_numpy(np.ascontiguousarray(opB)).to(dev)
    tC0 = torch.from_numpy(C0[:, :N].astype(np.float64)).to(dev)
    ref = alpha * (tA @ tB) + beta * tC0
    den = abs(alpha) * (tA.abs() @ tB.abs()) + abs(beta) * tC0.abs()
    err = (torch.from_numpy(C[:, :N].astype(np.float64)).to(dev) - ref).abs()
    bad = int((~(err <= (2 * (K + 2) * U32 + EXTRA[prec]) * den)).sum())
    assert bad == 0, (prec, ta, tb, M, N, K, alpha, beta, bad)
    # a sample of rows against the oracle (FP32, alpha = 1, beta = 0 cases)
    if prec == ls.FP32 and alpha == 1.0 and beta == 0.0:
        rows = np.unique(np.linspace(0, M - 1, num=min(M, 24)).astype(int))
        Asub = np.ascontiguousarray(opA[rows].astype(np.float32))
        Bm = np.ascontiguousarray(opB.astype(np.float32))
        Cref, absAB = oracle.sgemm_ref(0, 0, Asub, Bm, len(rows), N, K)
        Ke = np.full(len(rows) * N, float(K))
        nb, st = oracle.check_kv(C[rows, :N].reshape(-1), Cref.reshape(-1), absAB.reshape(-1), Ke, U32, 0.0)
        assert nb == 0, (ta, tb, M, N, K, st)


@pytest.mark.parametrize("ta,tb", TRANS)
def test_short_side_shapes(dev, ta, tb):
    for i, (M, N, K) in enumerate(SHAPES):
        assert is_short_side_plan(M, N, K), (M, N, K)
        run_general(dev, ta, tb, M, N, K, 1.0, 0.0, 0, 100 + i)


@pytest.mark.parametrize("ta,tb", TRANS)
def test_short_side_alpha_beta_padded(dev, ta, tb):
    for i, (M, N, K) in enumerate([(1500, 24, 640), (20, 2100, 900)]):
        assert is_short_side_plan(M, N, K), (M, N, K)
        for alpha, beta in [(0.5, 0.0), (-2.0, 1.5)]:
            run_general(dev, ta, tb, M, N, K, alpha, beta, 3, 200 + i)


@pytest.mark.parametrize("prec", [ls.BF16X3, ls.BF16, ls.BF16X6])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_short_side_split_modes(dev, prec, ta, tb):
    for i, (M, N, K) in enumerate([(1031, 5, 777), (4093, 31, 520), (7, 1500, 1000), (32, 2047, 512),
                                   (2048, 17, 1024)]):
        assert is_short_side_plan(M, N, K, prec), (prec, M, N, K)
        run_general(dev, ta, tb, M, N, K, 1.0, 0.0, 0, 300 + i, prec)
    run_general(dev, ta, tb, 1500, 24, 640, -2.0, 1.5, 3, 310, prec)


def test_short_side_reference_api_matches_general(dev):
    """LibShalom_sgemm_mp on a short-side shape equals sgemm_general bit for bit."""
    M, N, K = 4096, 16, 1024
    assert is_short_side_plan(M, N, K)
    r = oracle.Drand48(9)
    A = torch.from_numpy(r.matrix(M, K)).to(dev)
    B = torch.from_numpy(r.matrix(K, N)).to(dev)
    C1 = torch.full((M, N), float("nan"), device=dev)
    C2 = torch.full((M, N), float("nan"), device=dev)
    ls.sgemm_mp(0, 0, C1, A, B, M, N, K)
    ls.sgemm_general(0, 0, M, N, K, 1.0, A, K, B, N, 0.0, C2, N, ls.FP32)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
