"""A test-only torch.autograd.Function over the C ABI (ctypes): forward is
LibShalom_sgemm_ex, backward is the fused LibShalom_sgemm_backward (dA and dB
in one launch).  Its gradients are checked against

  * central finite differences in FP64 (step 1e-3) of the scalar loss
    L(A, B) = sum(G * op(A) op(B)) for a fixed upstream G, evaluated by the
    oracle's long-double GEMM (the loss is bilinear, so the central difference
    is exact up to rounding), at sampled entries; and
  * the gradient torch's autograd gives for torch.matmul in float64.

An im2col-convolution variant (weights shared by a batch of 32 images) uses a
batched plan forward and LibShalom_sgemm_backward_strided backward.  The
implicit-GEMM convolution (LibShalom_sconv2d_forward / _backward, on NCHW
images, no im2col matrix) is checked the same way: central differences of
sum(G * conv(X, W)) evaluated by the oracle's long-double convolution, and
torch's float64 autograd of conv2d.
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


class ShalomConv2d(torch.autograd.Function):
    """Y = conv2d(X, W) (NCHW, OIHW, stride 1) on the implicit-GEMM kernels
    (test-only): forward LibShalom_sconv2d_forward, backward
    LibShalom_sconv2d_backward (dX and dW in one launch)."""

    @staticmethod
    def forward(ctx, X, W, pad, prec):
        batch, C, H, Wd = X.shape
        Cout, _, R, S = W.shape
        P, Q = ls.conv2d_output_size(H, Wd, R, S, (1, 1), (pad, pad), (1, 1))
        Y = torch.empty(batch, Cout, P, Q, device=X.device, dtype=torch.float32)
        ls.sconv2d_forward(Y, W, X, batch, C, H, Wd, Cout, R, S, (1, 1), (pad, pad), (1, 1), prec)
        ctx.save_for_backward(X, W)
        ctx.dims = (pad, prec)
        return Y

    @staticmethod
    def backward(ctx, dY):
        X, W = ctx.saved_tensors
        pad, prec = ctx.dims
        batch, C, H, Wd = X.shape
        Cout, _, R, S = W.shape
        dX, dW = torch.empty_like(X), torch.empty_like(W)
        ls.sconv2d_backward(dX, dW, dY.contiguous(), W, X, batch, C, H, Wd, Cout, R, S, (1, 1), (pad, pad),
                            (1, 1), prec)
        return dX, dW, None, None


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
def test_implicit_conv_autograd_finite_differences(dev, prec):
    rng = np.random.default_rng(7)
    batch, C, H, Wd, Cout, R = 2, 3, 6, 7, 4, 3
    X = rng.standard_normal((batch, C, H, Wd)).astype(np.float32)
    W = (rng.standard_normal((Cout, C, R, R)) * np.sqrt(2.0 / (C * R * R))).astype(np.float32)
    G = rng.standard_normal((batch, Cout, H, Wd)).astype(np.float32)
    Xt = torch.tensor(X, device=dev, requires_grad=True)
    Wt = torch.tensor(W, device=dev, requires_grad=True)
    (ShalomConv2d.apply(Xt, Wt, 1, prec) * torch.tensor(G, device=dev)).sum().backward()
    gX, gW = Xt.grad.double().cpu().numpy(), Wt.grad.double().cpu().numpy()
    _, aX, _, aW = oracle.sconv2d_bwd_ref(X, W, G, (1, 1), (1, 1), (1, 1))

    def L(Xm, Wm):
        Y, _ = oracle.sconv2d_ref(Xm, Wm, (1, 1), (1, 1), (1, 1))
        return float(np.sum(Y * G.astype(np.float64)))

    # the loss is bilinear, so the central difference is exact up to rounding;
    # the perturbed operands are rounded to float32 by the oracle, so the
    # step is a power of two that float32 represents exactly around each entry
    h = 2.0 ** -10
    for _ in range(8):
        idx = tuple(int(rng.integers(n)) for n in X.shape)
        Xp, Xm = X.copy(), X.copy()
        Xp[idx] += h
        Xm[idx] -= h
        f = (L(Xp, W) - L(Xm, W)) / (Xp[idx] - Xm[idx])
        tol = (2 * Cout * R * R * U32 + EXTRA[prec]) * aX[idx] + 1e-6
        assert abs(gX[idx] - f) <= tol, ("X", idx, gX[idx], f)
    for _ in range(8):
        idx = tuple(int(rng.integers(n)) for n in W.shape)
        Wp, Wm = W.copy(), W.copy()
        Wp[idx] += h
        Wm[idx] -= h
        f = (L(X, Wp) - L(X, Wm)) / (Wp[idx] - Wm[idx])
        tol = (2 * batch * H * Wd * U32 + EXTRA[prec]) * aW[idx] + 1e-6
        assert abs(gW[idx] - f) <= tol, ("W", idx, gW[idx], f)


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
def test_implicit_conv_autograd_matches_torch_full_layer(dev, prec):
    """The first 3 x 3 layer at batch 32: gradients against torch's float64
    autograd of conv2d (computed on the CPU)."""
    batch, C, HW, Cout = 32, 64, 56, 64
    g = torch.Generator(device="cpu").manual_seed(23)
    X = torch.randn(batch, C, HW, HW, generator=g)
    W = torch.randn(Cout, C, 3, 3, generator=g) * (2.0 / (C * 9)) ** 0.5
    G = torch.randn(batch, Cout, HW, HW, generator=g)
    Xd, Wd = X.to(dev).requires_grad_(), W.to(dev).requires_grad_()
    (ShalomConv2d.apply(Xd, Wd, 1, prec) * G.to(dev)).sum().backward()
    X64, W64 = X.double().requires_grad_(), W.double().requires_grad_()
    (torch.nn.functional.conv2d(X64, W64, padding=1) * G.double()).sum().backward()
    Xa, Wa = X.double().abs().requires_grad_(), W.double().abs().requires_grad_()
    (torch.nn.functional.conv2d(Xa, Wa, padding=1) * G.double().abs()).sum().backward()
    tolX = (2 * Cout * 9 * U32 + EXTRA[prec]) * Xa.grad
    tolW = (2 * batch * HW * HW * U32 + EXTRA[prec]) * Wa.grad
    assert ((Xd.grad.double().cpu() - X64.grad).abs() <= tolX).all()
    assert ((Wd.grad.double().cpu() - W64.grad).abs() <= tolW).all()
