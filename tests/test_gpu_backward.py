"""Fused backward (one launch for dA and dB) on the GPU.

Reference: the FP64 products dA = dC op(B)^T, dB = op(A)^T dC (torch float64 on
the device for the large shapes -- FP64 is exact enough to serve as the
reference for FP32-level bounds), elementwise bound
    |g - g64| <= (2 * L * u + extra) * (|dC| |op(B)|^T)      (L = reduction length)
with u = 2^-24 and `extra` the split-precision term of the mode.  Split
reductions (slices summed inside the launch) and shared operands (gradient
summed over a batch) are covered, as are determinism and the reuse of the
in-launch arrival counters across calls.
"""
import os

import pytest
import torch

import libshalom_amd as ls
import oracle

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
EXTRA = {ls.FP32: 0.0, ls.BF16X6: 0.0, ls.BF16X3: 3 * 2.0 ** -16, ls.BF16: 2 * 2.0 ** -8}
TRANS = [(0, 0), (0, 1), (1, 0), (1, 1)]


def ref_grads(ta, tb, A, B, dC):
    """FP64 gradients (stored like A and B) and their |.|-products for the bound."""
    A64, B64, dC64 = A.double(), B.double(), dC.double()
    opA = A64.transpose(-1, -2) if ta else A64
    opB = B64.transpose(-1, -2) if tb else B64
    gA = dC64 @ opB.transpose(-1, -2).contiguous()
    gB = opA.transpose(-1, -2).contiguous() @ dC64
    bA = dC64.abs() @ opB.abs().transpose(-1, -2).contiguous()
    bB = opA.abs().transpose(-1, -2).contiguous() @ dC64.abs()
    if ta:
        gA, bA = gA.transpose(-1, -2), bA.transpose(-1, -2)
    if tb:
        gB, bB = gB.transpose(-1, -2), bB.transpose(-1, -2)
    return gA, gB, bA, bB


def assert_within(got, ref, absprod, L, prec, what):
    bound = (2 * max(L, 1) * U32 + EXTRA[prec]) * absprod
    err = (got.double() - ref).abs()
    bad = (err > bound) | ~torch.isfinite(got.double())
    assert not bad.any(), f"{what}: {int(bad.sum())} elements out of bound, max err {float(err.max()):.3g}"


def run_single(dev, ta, tb, M, N, K, prec, seed=0):
    g = torch.Generator(device="cpu").manual_seed(seed)
    A = torch.randn(*oracle.a_shape(ta, M, K), generator=g).float().to(dev)
    B = torch.randn(*oracle.b_shape(tb, N, K), generator=g).float().to(dev)
    dC = torch.randn(M, N, generator=g).float().to(dev)
    dA = torch.full_like(A, float("nan"))
    dB = torch.full_like(B, float("nan"))
    ls.sgemm_backward(ta, tb, dC, A, B, dA, dB, M, N, K, prec)
    torch.cuda.synchronize()
    return A, B, dC, dA, dB


# (M, N, K): small and ragged; dA reduction long enough to be sliced; dB
# reduction sliced; every output dimension >= 128 (128 x 128 tiles).
SHAPES = [(17, 33, 5), (1, 1, 1), (64, 8192, 96), (2048, 48, 64), (300, 200, 150), (129, 257, 130)]


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_fused_backward_single(dev, ta, tb, prec):
    for (M, N, K) in SHAPES:
        print(f"backward ta={ta} tb={tb} prec={prec} M={M} N={N} K={K}", flush=True)
        A, B, dC, dA, dB = run_single(dev, ta, tb, M, N, K, prec, seed=M + N + K)
        gA, gB, bA, bB = ref_grads(ta, tb, A, B, dC)
        assert_within(dA, gA, bA, N, prec, f"dA {M}x{N}x{K}")
        assert_within(dB, gB, bB, M, prec, f"dB {M}x{N}x{K}")


@pytest.mark.parametrize("prec", [ls.BF16, ls.BF16X6])
def test_fused_backward_other_modes(dev, prec):
    for (M, N, K) in [(64, 8192, 96), (300, 200, 150)]:
        A, B, dC, dA, dB = run_single(dev, 0, 0, M, N, K, prec, seed=7)
        gA, gB, bA, bB = ref_grads(0, 0, A, B, dC)
        assert_within(dA, gA, bA, N, prec, "dA")
        assert_within(dB, gB, bB, M, prec, "dB")


def test_backward_empty_reductions(dev):
    # N = 0: dC is M x 0, so dA = 0 (dB is K x 0); M = 0: dB = 0
    for (M, N, K) in [(5, 0, 7), (0, 6, 3)]:
        A = torch.randn(M, K, device=dev)
        B = torch.randn(K, N, device=dev)
        dC = torch.randn(M, N, device=dev)
        dA = torch.full_like(A, float("nan"))
        dB = torch.full_like(B, float("nan"))
        ls.sgemm_backward(0, 0, dC, A, B, dA, dB, M, N, K)
        torch.cuda.synchronize()
        assert (dA == 0).all() and (dB == 0).all()


def make_conv(dev, Cout, HW, Kc, batch, seed, gen_dtype=torch.float32):
    """im2col convolution GEMMs: C_i = W X_i with W (Cout x Kc) ~ N(0, 2/Kc),
    X_i (Kc x HW) ~ N(0, 1), upstream dC_i ~ N(0, 1)."""
    g = torch.Generator(device=dev).manual_seed(seed)
    W = torch.randn(Cout, Kc, generator=g, device=dev, dtype=gen_dtype) * (2.0 / Kc) ** 0.5
    X = torch.randn(batch, Kc, HW, generator=g, device=dev, dtype=gen_dtype)
    dC = torch.randn(batch, Cout, HW, generator=g, device=dev, dtype=gen_dtype)
    return W.float().contiguous(), X.float().contiguous(), dC.float().contiguous()


def conv_backward(W, X, dC, prec):
    batch, Kc, HW = X.shape
    Cout = W.shape[0]
    dW = torch.full_like(W, float("nan"))
    dX = torch.full_like(X, float("nan"))
    ls.sgemm_backward_strided(0, 0, dC, W, X, dW, dX, Cout, HW, Kc, batch, 0, Kc * HW, Cout * HW, prec)
    torch.cuda.synchronize()
    return dW, dX


def check_conv(W, X, dC, dW, dX, prec):
    batch = X.shape[0]
    W64, X64, dC64 = W.double(), X.double(), dC.double()
    gW = torch.zeros_like(W64)
    bW = torch.zeros_like(W64)
    for i in range(batch):  # dW = sum_i dC_i X_i^T
        gW += dC64[i] @ X64[i].T.contiguous()
        bW += dC64[i].abs() @ X64[i].abs().T.contiguous()
    gX = W64.T.contiguous() @ dC64  # dX_i = W^T dC_i
    bX = W64.abs().T.contiguous() @ dC64.abs()
    assert_within(dW, gW, bW, batch * X.shape[2], prec, "dW")
    assert_within(dX, gX, bX, W.shape[0], prec, "dX")


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
def test_strided_conv_small(dev, prec):
    for (Cout, HW, Kc, batch) in [(16, 196, 72, 4), (8, 49, 36, 3), (40, 100, 300, 5)]:
        W, X, dC = make_conv(dev, Cout, HW, Kc, batch, seed=Cout)
        dW, dX = conv_backward(W, X, dC, prec)
        check_conv(W, X, dC, dW, dX, prec)


# The four im2col layers, batch 32: (M = Cout, N = H*W, K = Cin*9).
CONV_SHAPES = [(64, 3136, 576), (128, 784, 1152), (256, 196, 2304), (512, 49, 4608)]


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
@pytest.mark.parametrize("shape", CONV_SHAPES)
def test_conv_backward_full_shapes(dev, shape, prec):
    Cout, HW, Kc = shape
    print(f"conv backward {shape} batch 32 prec={prec}", flush=True)
    W, X, dC = make_conv(dev, Cout, HW, Kc, 32, seed=11)
    dW, dX = conv_backward(W, X, dC, prec)
    check_conv(W, X, dC, dW, dX, prec)


def test_strided_shared_b_and_unshared(dev):
    # B shared (its gradient summed over the batch), A per image; and both per image
    batch, M, N, K = 6, 70, 90, 50
    g = torch.Generator(device="cpu").manual_seed(5)
    A = torch.randn(batch, M, K, generator=g).to(dev)
    B = torch.randn(K, N, generator=g).to(dev)
    dC = torch.randn(batch, M, N, generator=g).to(dev)
    dA = torch.full_like(A, float("nan"))
    dB = torch.full_like(B, float("nan"))
    ls.sgemm_backward_strided(0, 0, dC, A, B, dA, dB, M, N, K, batch, M * K, 0, M * N)
    Bb = torch.stack([B] * batch)
    dA2 = torch.full_like(A, float("nan"))
    dB2 = torch.full_like(Bb, float("nan"))
    ls.sgemm_backward_strided(0, 0, dC, A, Bb, dA2, dB2, M, N, K, batch, M * K, K * N, M * N)
    torch.cuda.synchronize()
    gA, gB, bA, bB = ref_grads(0, 0, A, Bb, dC)
    assert_within(dA, gA, bA, N, ls.FP32, "dA (B shared)")
    assert_within(dA2, gA, bA, N, ls.FP32, "dA (per image)")
    assert_within(dB2, gB, bB, M, ls.FP32, "dB (per image)")
    assert_within(dB, gB.sum(0), bB.sum(0), batch * M, ls.FP32, "dB (summed)")


def test_backward_deterministic_and_counter_reuse(dev):
    # two sliced shapes alternately on one stream: the arrival counters are
    # reset by every launch, and the slice order fixes the result bits
    first = {}
    for rep in range(3):
        for (M, N, K) in [(64, 8192, 96), (2048, 48, 64)]:
            A, B, dC, dA, dB = run_single(dev, 0, 1, M, N, K, ls.BF16X3, seed=1)
            if rep == 0:
                gA, gB, bA, bB = ref_grads(0, 1, A, B, dC)
                assert_within(dA, gA, bA, N, ls.BF16X3, "dA")
                assert_within(dB, gB, bB, M, ls.BF16X3, "dB")
                first[(M, N, K)] = (dA.clone(), dB.clone())
            else:
                assert torch.equal(dA, first[(M, N, K)][0]) and torch.equal(dB, first[(M, N, K)][1])


def test_backward_workspace_failure(dev):
    M, N, K = 64, 8192, 96  # sliced: needs the fixup workspace
    A = torch.randn(M, K, device=dev)
    B = torch.randn(K, N, device=dev)
    dC = torch.randn(M, N, device=dev)
    dA, dB = torch.empty_like(A), torch.empty_like(B)
    s = torch.cuda.Stream()  # a stream with no workspace yet
    os.environ["LIBSHALOM_FAIL_WORKSPACE"] = "1"
    try:
        with pytest.raises(ls.LibShalomError) as e:
            ls.sgemm_backward(0, 0, dC, A, B, dA, dB, M, N, K, ls.FP32, stream=s)
        assert e.value.status == 4
    finally:
        del os.environ["LIBSHALOM_FAIL_WORKSPACE"]
    torch.cuda.synchronize()


def strided_ref(ta, tb, A, B):
    A64, B64 = A.double(), B.double()
    opA = A64.transpose(-1, -2).contiguous() if ta else A64
    opB = B64.transpose(-1, -2).contiguous() if tb else B64
    return opA @ opB, opA.abs() @ opB.abs()


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_strided_batched_forward(dev, ta, tb, prec):
    g = torch.Generator(device="cpu").manual_seed(17)
    for (M, N, K, batch, shareA, shareB) in [(70, 90, 300, 5, False, False), (64, 3136, 576, 3, True, False),
                                            (512, 49, 4608, 2, True, False), (33, 17, 5000, 4, False, True),
                                            (1, 1, 1, 3, False, False), (200, 150, 130, 2, False, False)]:
        print(f"strided ta={ta} tb={tb} prec={prec} {M}x{N}x{K} batch {batch}", flush=True)
        A = torch.randn(*((1,) if shareA else (batch,)), *oracle.a_shape(ta, M, K), generator=g).to(dev)
        B = torch.randn(*((1,) if shareB else (batch,)), *oracle.b_shape(tb, N, K), generator=g).to(dev)
        C = torch.full((batch, M, N), float("nan"), device=dev)
        ls.sgemm_strided_batched(ta, tb, C, A, B, M, N, K, batch, 0 if shareA else M * K, 0 if shareB else K * N,
                                 M * N, prec)
        torch.cuda.synchronize()
        ref, absp = strided_ref(ta, tb, A.expand(batch, *A.shape[1:]), B.expand(batch, *B.shape[1:]))
        assert_within(C, ref, absp, K, prec, f"C {M}x{N}x{K}")


def test_strided_batched_sum_over_batch(dev):
    # strideC = 0: C = sum_i A_i B_i (one output, reduction over images and K)
    batch, M, N, K = 7, 96, 80, 200
    A = torch.randn(batch, M, K, device=dev)
    B = torch.randn(batch, K, N, device=dev)
    C = torch.full((M, N), float("nan"), device=dev)
    ls.sgemm_strided_batched(0, 0, C, A, B, M, N, K, batch, M * K, K * N, 0)
    torch.cuda.synchronize()
    ref, absp = strided_ref(0, 0, A, B)
    assert_within(C, ref.sum(0), absp.sum(0), batch * K, ls.FP32, "sum")
