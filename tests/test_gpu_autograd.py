"""A test-only torch.autograd.Function over the C ABI (ctypes): forward is
LibShalom_sgemm_ex, backward is the fused LibShalom_sgemm_backward (dA and dB
in one launch).  Its gradients are checked against

  * central finite differences in FP64 (step 1e-3) of the scalar loss
    L(A, B) = sum(G * op(A) op(B)) for a fixed upstream G, evaluated by the
    oracle's long-double GEMM (the loss is bilinear, so the central difference
    is exact up to rounding), at sampled entries; and
  * the gradient torch's autograd gives for torch.matmul in float64.

An im2col-convolution variant (weights shared by a batch of 32 images) uses a
batched plan forward and LibShalom_sgemm_backward_strided backward.
"""
import numpy as np
import pytest
import torch

import libshalom_amd as ls
import oracle

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
EXTRA = {ls.FP32: 0.0, ls.BF16X3: 3 * 2.0 ** -16}


class ShalomMatmul(torch.autograd.Function):
    """C = op(A) op(B) on the library's kernels (test-only)."""

    @staticmethod
    def forward(ctx, A, B, ta, tb, prec):
        M = A.shape[1] if ta else A.shape[0]
        K = A.shape[0] if ta else A.shape[1]
        N = B.shape[0] if tb else B.shape[1]
        C = torch.empty(M, N, device=A.device, dtype=torch.float32)
        ls.sgemm_ex(ta, tb, C, A, B, M, N, K, prec)
        ctx.save_for_backward(A, B)
        ctx.dims = (ta, tb, prec, M, N, K)
        return C

    @staticmethod
    def backward(ctx, dC):
        A, B = ctx.saved_tensors
        ta, tb, prec, M, N, K = ctx.dims
        dA, dB = torch.empty_like(A), torch.empty_like(B)
        ls.sgemm_backward(ta, tb, dC.contiguous(), A, B, dA, dB, M, N, K, prec)
        return dA, dB, None, None, None


def loss64(A, B, G, ta, tb):
    """sum(G * op(A) op(B)) with the oracle's long-double GEMM."""
    M, N = G.shape
    K = A.shape[0] if ta else A.shape[1]
    C, _ = oracle.dgemm_ref(ta, tb, A, B, M, N, K)
    return float(np.sum(G * C.reshape(M, N)))


def fd_entries(X, which, A, B, G, ta, tb, idx, h=1e-3):
    out = []
    for (i, j) in idx:
        Xp, Xm = X.copy(), X.copy()
        Xp[i, j] += h
        Xm[i, j] -= h
        if which == "A":
            lp, lm = loss64(Xp, B, G, ta, tb), loss64(Xm, B, G, ta, tb)
        else:
            lp, lm = loss64(A, Xp, G, ta, tb), loss64(A, Xm, G, ta, tb)
        out.append((lp - lm) / (2 * h))
    return np.array(out)


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_autograd_vs_finite_differences_and_matmul(dev, ta, tb, prec):
    rng = np.random.default_rng(10 * ta + tb)
    for (M, N, K) in [(24, 40, 16), (70, 33, 129)]:
        A = rng.standard_normal(oracle.a_shape(ta, M, K))
        B = rng.standard_normal(oracle.b_shape(tb, N, K))
        G = rng.standard_normal((M, N))
        At = torch.tensor(A, dtype=torch.float32, device=dev, requires_grad=True)
        Bt = torch.tensor(B, dtype=torch.float32, device=dev, requires_grad=True)
        Gt = torch.tensor(G, dtype=torch.float32, device=dev)
        C = ShalomMatmul.apply(At, Bt, ta, tb, prec)
        (C * Gt).sum().backward()
        gA = At.grad.double().cpu().numpy()
        gB = Bt.grad.double().cpu().numpy()
        # the inputs the kernels saw (FP32), in FP64 for the references
        A32 = At.detach().double().cpu().numpy()
        B32 = Bt.detach().double().cpu().numpy()
        G32 = Gt.double().cpu().numpy()
        # bounds: a GEMM over the reduction of length N (dA) or M (dB)
        opA = A32.T if ta else A32
        opB = B32.T if tb else B32
        bA = np.abs(G32) @ np.abs(opB).T
        bB = np.abs(opA).T @ np.abs(G32)
        bA, bB = (bA.T if ta else bA), (bB.T if tb else bB)
        tolA = (2 * N * U32 + EXTRA[prec]) * bA + 1e-9
        tolB = (2 * M * U32 + EXTRA[prec]) * bB + 1e-9
        # central finite differences at sampled entries
        ia = [(int(rng.integers(A.shape[0])), int(rng.integers(A.shape[1]))) for _ in range(12)]
        ib = [(int(rng.integers(B.shape[0])), int(rng.integers(B.shape[1]))) for _ in range(12)]
        fdA = fd_entries(A32, "A", A32, B32, G32, ta, tb, ia)
        fdB = fd_entries(B32, "B", A32, B32, G32, ta, tb, ib)
        # FD rounding: |L| * 2^-60 / h, far below the FP32 bound
        for (i, j), f in zip(ia, fdA):
            assert abs(gA[i, j] - f) <= tolA[i, j] + 1e-8, (M, N, K, "A", i, j, gA[i, j], f)
        for (i, j), f in zip(ib, fdB):
            assert abs(gB[i, j] - f) <= tolB[i, j] + 1e-8, (M, N, K, "B", i, j, gB[i, j], f)
        # matmul reference gradient (torch autograd, float64)
        A64 = torch.tensor(A32, dtype=torch.float64, requires_grad=True)
        B64 = torch.tensor(B32, dtype=torch.float64, requires_grad=True)
        oA = A64.T if ta else A64
        oB = B64.T if tb else B64
        (torch.matmul(oA, oB) * torch.tensor(G32)).sum().backward()
        assert np.all(np.abs(gA - A64.grad.numpy()) <= tolA)
        assert np.all(np.abs(gB - B64.grad.numpy()) <= tolB)


class ShalomConvGemm(torch.autograd.Function):
    """Y_i = W X_i for a batch of im2col matrices X (batch x Kc x HW), W shared
    (test-only): forward through a batched plan, backward through the strided
    fused backward with W's stride 0 (dW summed over the batch)."""

    @staticmethod
    def forward(ctx, W, X, prec):
        batch, Kc, HW = X.shape
        Cout = W.shape[0]
        Y = torch.empty(batch, Cout, HW, device=X.device, dtype=torch.float32)
        plan = ls.BatchedPlan(0, 0, [Y[i] for i in range(batch)], [W] * batch, [X[i] for i in range(batch)],
                              [Cout] * batch, [HW] * batch, [Kc] * batch, prec)
        plan.execute()
        torch.cuda.synchronize()
        plan.close()
        ctx.save_for_backward(W, X)
        ctx.prec = prec
        return Y

    @staticmethod
    def backward(ctx, dY):
        W, X = ctx.saved_tensors
        batch, Kc, HW = X.shape
        Cout = W.shape[0]
        dW, dX = torch.empty_like(W), torch.empty_like(X)
        ls.sgemm_backward_strided(0, 0, dY.contiguous(), W, X, dW, dX, Cout, HW, Kc, batch, 0, Kc * HW,
                                  Cout * HW, ctx.prec)
        return dW, dX, None


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
def test_conv_autograd_finite_differences(dev, prec):
    # a reduced layer (batch 3) so that each FD loss evaluation is cheap
    rng = np.random.default_rng(3)
    Cout, HW, Kc, batch = 12, 30, 27, 3
    W = rng.standard_normal((Cout, Kc)) * np.sqrt(2.0 / Kc)
    X = rng.standard_normal((batch, Kc, HW))
    G = rng.standard_normal((batch, Cout, HW))
    Wt = torch.tensor(W, dtype=torch.float32, device=dev, requires_grad=True)
    Xt = torch.tensor(X, dtype=torch.float32, device=dev, requires_grad=True)
    Y = ShalomConvGemm.apply(Wt, Xt, prec)
    (Y * torch.tensor(G, dtype=torch.float32, device=dev)).sum().backward()
    gW = Wt.grad.double().cpu().numpy()
    gX = Xt.grad.double().cpu().numpy()
    W32 = Wt.detach().double().cpu().numpy()
    X32 = Xt.detach().double().cpu().numpy()
    G32 = torch.tensor(G, dtype=torch.float32).double().numpy()

    def L(Wm, Xm):
        return sum(loss64(Wm, Xm[i], G32[i], 0, 0) for i in range(batch))

    h = 1e-3
    bW = sum(np.abs(G32[i]) @ np.abs(X32[i]).T for i in range(batch))
    tolW = (2 * batch * HW * U32 + EXTRA[prec]) * bW + 1e-8
    for _ in range(8):
        i, j = int(rng.integers(Cout)), int(rng.integers(Kc))
        Wp, Wm = W32.copy(), W32.copy()
        Wp[i, j] += h
        Wm[i, j] -= h
        f = (L(Wp, X32) - L(Wm, X32)) / (2 * h)
        assert abs(gW[i, j] - f) <= tolW[i, j], ("W", i, j, gW[i, j], f)
    for _ in range(8):
        b, i, j = int(rng.integers(batch)), int(rng.integers(Kc)), int(rng.integers(HW))
        bX = np.abs(W32).T @ np.abs(G32[b])
        tolX = (2 * Cout * U32 + EXTRA[prec]) * bX[i, j] + 1e-8
        Xp, Xm = X32.copy(), X32.copy()
        Xp[b, i, j] += h
        Xm[b, i, j] -= h
        f = (L(W32, Xp) - L(W32, Xm)) / (2 * h)
        assert abs(gX[b, i, j] - f) <= tolX, ("X", b, i, j, gX[b, i, j], f)


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
def test_conv_autograd_matmul_reference_full_batch(dev, prec):
    # first layer shape, batch 32: gradients vs torch autograd of matmul in float64
    Cout, HW, Kc, batch = 64, 3136, 576, 32
    g = torch.Generator(device=dev).manual_seed(21)
    W = (torch.randn(Cout, Kc, generator=g, device=dev) * (2.0 / Kc) ** 0.5).requires_grad_()
    X = torch.randn(batch, Kc, HW, generator=g, device=dev).requires_grad_()
    G = torch.randn(batch, Cout, HW, generator=g, device=dev)
    (ShalomConvGemm.apply(W, X, prec) * G).sum().backward()
    W64 = W.detach().double().requires_grad_()
    X64 = X.detach().double().requires_grad_()
    (torch.matmul(W64, X64) * G.double()).sum().backward()
    bW = (G.double().abs() @ X64.detach().abs().transpose(1, 2).contiguous()).sum(0)
    bX = W64.detach().abs().T.contiguous() @ G.double().abs()
    tolW = (2 * batch * HW * U32 + EXTRA[prec]) * bW
    tolX = (2 * Cout * U32 + EXTRA[prec]) * bX
    assert ((W.grad.double() - W64.grad).abs() <= tolW).all()
    assert ((X.grad.double() - X64.grad).abs() <= tolX).all()
