"""The allocation-free entry point (LibShalom_sgemm_general_noalloc) against
LibShalom_sgemm_general, through the C-ABI: two calls in a row, the second
on a smaller problem, both with K split across slices (so the static scratch
array's counters and partial tiles are used and must come back clean), each
checked against the oracle and bit for bit against the allocating entry
point; then alpha / beta on a padded C."""
import numpy as np
import pytest
import torch

import libshalom_amd as ls
import oracle

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


def run_pair(dev, M, N, K, tb, prec, seed):
    r = oracle.Drand48(seed)
    A = r.matrix(M, K)
    B = r.matrix(*oracle.b_shape(tb, N, K))
    dA = torch.from_numpy(A).to(dev)
    dB = torch.from_numpy(B).to(dev)
    ldb = K if tb else N
    C1 = torch.full((M, N), float("nan"), device=dev)
    C2 = torch.full((M, N), float("nan"), device=dev)
    ls.sgemm_general_noalloc(0, tb, M, N, K, 1.0, dA, K, dB, ldb, 0.0, C1, N, prec)
    ls.sgemm_general(0, tb, M, N, K, 1.0, dA, K, dB, ldb, 0.0, C2, N, prec)
    torch.cuda.synchronize()
    c1 = C1.cpu().numpy()
    Cref, absAB = oracle.sgemm_ref(0, tb, A, B, M, N, K)
    bad, st = oracle.check(c1, Cref, absAB, K, oracle.U_F32)
    assert bad == 0, (M, N, K, st)
    assert np.array_equal(c1, C2.cpu().numpy()), (M, N, K)


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X6])
def test_noalloc_matches_allocating_entry(dev, prec):
    # both plans split K (8 and 4 slices); the second call is smaller.  (M
    # below 1024 keeps them off the FP32 short-side kernel, which does not
    # split K and uses no scratch.)
    for (M, N, K, tb, seed) in [(1000, 16, 1024, 0, 1), (512, 32, 512, 1, 2)]:
        plan = ls.query_plan(prec, M, N, K)
        assert plan["ksplit"] > 1, plan
        run_pair(dev, M, N, K, tb, prec, seed)


def test_noalloc_alpha_beta_padded(dev):
    M, N, K, ldc = 2048, 24, 768, 40
    r = oracle.Drand48(5)
    A = r.matrix(M, K)
    B = r.matrix(K, N)
    C0 = r.matrix(M, ldc)
    dA = torch.from_numpy(A).to(dev)
    dB = torch.from_numpy(B).to(dev)
    C1 = torch.from_numpy(C0.copy()).to(dev)
    C2 = torch.from_numpy(C0.copy()).to(dev)
    ls.sgemm_general_noalloc(0, 0, M, N, K, 0.5, dA, K, dB, N, -1.5, C1, ldc)
    ls.sgemm_general(0, 0, M, N, K, 0.5, dA, K, dB, N, -1.5, C2, ldc)
    torch.cuda.synchronize()
    c1 = C1.cpu().numpy()
    assert np.array_equal(c1, C2.cpu().numpy())
    assert np.array_equal(c1[:, N:], C0[:, N:])  # padding untouched
    Cref, absAB = oracle.sgemm_ref(0, 0, A, B, M, N, K)
    ref = 0.5 * Cref + -1.5 * C0[:, :N].astype(np.float64)
    bound = 2 * (K + 2) * U32 * (0.5 * absAB + 1.5 * np.abs(C0[:, :N]))
    assert (np.abs(c1[:, :N] - ref) <= bound).all()
