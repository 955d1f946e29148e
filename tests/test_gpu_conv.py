"""Convolution forward as an implicit GEMM on the GPU (LibShalom_sconv2d_forward).

The im2col matrix is never stored: the B-operand loads gather it from the
NCHW input.  References: the oracle's long-double convolution (small
geometries: stride, padding, dilation, non-square filters, reductions cut
into slices of whole channels) and torch's float64 conv2d on the CPU (the
four 3 x 3 layers whose im2col GEMMs the reference's CNN benchmark times,
benchmark/CNN_kernels/LibShalom_sgemm_mp.c:86-116, and a case that takes the
128 x 128 tiles).  Elementwise bound
    |y - y64| <= (2 * K * u + extra) * (|W| |X|)_i,     K = C * R * S,
u = 2^-24 and `extra` the split-precision term of the mode.  Also: results
bit-identical to the explicit path (torch's unfold, then the strided batched
GEMM) when neither cuts the reduction, deterministic, and no write outside Y.
"""
import numpy as np
import pytest
import torch

import libshalom_amd as ls
import oracle

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
EXTRA = {ls.FP32: 0.0, ls.BF16X6: 0.0, ls.BF16X3: 3 * 2.0 ** -16, ls.BF16: 2 * 2.0 ** -8}
PRECS = [ls.FP32, ls.BF16X3, ls.BF16, ls.BF16X6]

# (batch, C, H, W, Cout, R, S, stride, padding, dilation)
GEOMS = [
    (2, 3, 9, 11, 5, 3, 3, (1, 1), (1, 1), (1, 1)),
    (1, 4, 8, 8, 6, 1, 1, (1, 1), (0, 0), (1, 1)),
    (2, 5, 13, 7, 6, 3, 2, (2, 1), (1, 0), (1, 1)),
    (1, 3, 15, 15, 4, 3, 3, (1, 1), (2, 2), (2, 2)),
    (3, 2, 6, 5, 3, 5, 5, (1, 1), (3, 3), (1, 1)),
    (1, 2, 5, 9, 2, 2, 4, (3, 2), (0, 1), (1, 2)),
    (2, 16, 12, 12, 70, 3, 3, (1, 1), (1, 1), (1, 1)),   # 70 output channels: two row tiles
    (1, 40, 10, 10, 33, 3, 3, (1, 1), (1, 1), (1, 1)),   # K = 360: reduction cut into slices
    (1, 64, 7, 7, 32, 3, 3, (1, 1), (1, 1), (1, 1)),     # K = 576
    (2, 7, 19, 23, 9, 7, 7, (2, 2), (3, 3), (1, 1)),     # 7 x 7, stride 2 (R*S = 49)
    (1, 1, 130, 3, 2, 3, 3, (1, 1), (1, 1), (1, 1)),     # a single input channel, tall image
    (2, 9, 9, 10, 7, 4, 8, (1, 1), (2, 4), (1, 1)),      # R*S = 32: the last size with a tap mask
    (1, 5, 8, 15, 6, 3, 11, (1, 1), (1, 5), (1, 1)),     # R*S = 33: range compares instead
    (1, 300, 5, 5, 20, 1, 1, (1, 1), (0, 0), (1, 1)),    # 1 x 1, K = 300 cut into slices
    (2, 4, 17, 13, 5, 3, 3, (2, 2), (2, 1), (2, 1)),     # stride and dilation together
    (1, 6, 12, 12, 9, 4, 4, (2, 2), (1, 1), (1, 1)),     # 4 x 4, stride 2
    (1, 3, 11, 10, 4, 3, 3, (3, 3), (1, 0), (1, 1)),     # stride 3
    (1, 3, 16, 16, 4, 5, 5, (4, 2), (2, 2), (1, 1)),     # stride (4, 2), 5 x 5
    (2, 3, 14, 13, 5, 6, 6, (3, 2), (2, 1), (1, 2)),     # R*S = 36, stride (3, 2), dilation (1, 2)
    (2, 2, 23, 21, 3, 7, 7, (2, 2), (3, 3), (1, 1)),     # the 7 x 7 stride-2 stem, odd sizes
    # Q a multiple of 4 and stride 1 along W: dW's gather in 16-byte runs
    (2, 3, 10, 12, 4, 3, 7, (1, 1), (1, 3), (1, 1)),     # up to 3 leading padding columns per run
    (1, 2, 9, 16, 3, 2, 3, (1, 1), (0, 2), (1, 2)),      # dilation 2 along W
    (2, 3, 11, 8, 5, 3, 3, (2, 1), (1, 1), (1, 1)),      # stride 2 along H only
]


def _inputs(batch, C, H, W, Cout, R, S, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (batch, C, H, W)).astype(np.float32)
    Wt = rng.uniform(-1, 1, (Cout, C, R, S)).astype(np.float32)
    return X, Wt


def _run(X, Wt, st, pad, dil, prec, guard=0):
    """Convolution on the device; Y is followed by `guard` NaN sentinels."""
    batch, C, H, W = X.shape
    Cout, _, R, S = Wt.shape
    P, Q = ls.conv2d_output_size(H, W, R, S, st, pad, dil)
    n = batch * Cout * P * Q
    buf = torch.full((n + guard,), float("nan"), device="cuda")
    Xd = torch.from_numpy(X).cuda()
    Wd = torch.from_numpy(Wt).cuda()
    ls.sconv2d_forward(buf, Wd, Xd, batch, C, H, W, Cout, R, S, st, pad, dil, prec)
    torch.cuda.synchronize()
    return buf[:n].reshape(batch, Cout, P, Q).cpu().numpy(), buf[n:].cpu().numpy()


def _assert_within(y, ref, absprod, K, prec, what):
    bound = (2 * max(K, 1) * U32 + EXTRA[prec]) * absprod
    err = np.abs(y.astype(np.float64) - ref)
    bad = (err > bound) | ~np.isfinite(y)
    assert not bad.any(), f"{what}: {int(bad.sum())} of {y.size} out of bound, max err {float(err.max()):.3g}"


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("g", GEOMS)
def test_conv_matches_oracle(g, prec):
    batch, C, H, W, Cout, R, S, st, pad, dil = g
    print(f"conv {g} precision {prec}", flush=True)
    X, Wt = _inputs(batch, C, H, W, Cout, R, S, seed=sum(g[:7]))
    y, tail = _run(X, Wt, st, pad, dil, prec, guard=64)
    ref, absprod = oracle.sconv2d_ref(X, Wt, st, pad, dil)
    assert y.shape == ref.shape
    _assert_within(y, ref, absprod, C * R * S, prec, f"conv {g}")
    assert np.isnan(tail).all(), "write past the end of Y"


def _torch_ref(X, Wt, st, pad, dil):
    x, w = torch.from_numpy(X).double(), torch.from_numpy(Wt).double()
    ref = torch.nn.functional.conv2d(x, w, stride=st, padding=pad, dilation=dil).numpy()
    absprod = torch.nn.functional.conv2d(x.abs(), w.abs(), stride=st, padding=pad, dilation=dil).numpy()
    return ref, absprod


# The 3 x 3 layers (Cout, Cin, H = W), batch 4, and one that fills 128 x 128 tiles.
LAYERS = [(64, 64, 56, 4), (128, 128, 28, 4), (256, 256, 14, 4), (512, 512, 7, 4), (256, 16, 32, 16)]


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
@pytest.mark.parametrize("layer", LAYERS)
def test_conv_layers_match_fp64(layer, prec):
    Cout, C, HW, batch = layer
    print(f"conv layer {layer} precision {prec}", flush=True)
    X, Wt = _inputs(batch, C, HW, HW, Cout, 3, 3, seed=Cout + C)
    y, _ = _run(X, Wt, (1, 1), (1, 1), (1, 1), prec)
    ref, absprod = _torch_ref(X, Wt, (1, 1), (1, 1), (1, 1))
    _assert_within(y, ref, absprod, C * 9, prec, f"layer {layer}")


@pytest.mark.parametrize("prec", PRECS)
def test_conv_bit_identical_to_explicit_im2col(prec):
    """Gathering the im2col matrix in the operand loads feeds the tile
    pipeline the same values as loading it from memory (padding taps are
    zeros either way), so with the same tile and an uncut reduction
    (K = 144) the sums are the same, bit for bit."""
    batch, C, H, W, Cout, R, S = 3, 16, 20, 20, 70, 3, 3
    print(f"explicit vs implicit, precision {prec}", flush=True)
    X, Wt = _inputs(batch, C, H, W, Cout, R, S, seed=5)
    y, _ = _run(X, Wt, (1, 1), (1, 1), (1, 1), prec)
    Xd = torch.from_numpy(X).cuda()
    Wd = torch.from_numpy(Wt).cuda()
    cols = torch.nn.functional.unfold(Xd, (R, S), padding=1).contiguous()
    K, PQ = C * R * S, H * W
    assert cols.shape == (batch, K, PQ)
    y2 = torch.full((batch, Cout, PQ), float("nan"), device="cuda")
    ls.sgemm_strided_batched(0, 0, y2, Wd, cols, Cout, PQ, K, batch, 0, K * PQ, Cout * PQ, prec)
    torch.cuda.synchronize()
    assert np.array_equal(y.reshape(batch, Cout, PQ), y2.cpu().numpy())


def test_conv_deterministic_and_sliced_sum_in_order():
    """A sliced reduction (K = 2304, slices summed inside the launch in
    slice order) gives the same bits on every call."""
    X, Wt = _inputs(2, 256, 14, 14, 64, 3, 3, seed=9)
    y1, _ = _run(X, Wt, (1, 1), (1, 1), (1, 1), ls.BF16X3)
    y2, _ = _run(X, Wt, (1, 1), (1, 1), (1, 1), ls.BF16X3)
    assert np.array_equal(y1, y2)


def test_conv_empty_batch_is_a_no_op():
    y = torch.full((8,), 3.0, device="cuda")
    ls.sconv2d_forward(y, y, y, 0, 3, 8, 8, 4, 3, 3, (1, 1), (1, 1), (1, 1), ls.FP32)
    torch.cuda.synchronize()
    assert (y.cpu() == 3.0).all()


# Backward: stride-1 geometries, and strided ones (dX through the fractionally
# strided gather: the per-tile tap mask for filters of at most 32 taps, the
# per-element remainder test beyond).
BWD_GEOMS = [g for g in GEOMS if g[7] == (1, 1)]
BWD_STRIDED = [g for g in GEOMS if g[7] != (1, 1)]


def _run_bwd(X, Wt, dY, st, pad, dil, prec, want_dx=True, guard=64):
    batch, C, H, W = X.shape
    Cout, _, R, S = Wt.shape
    nx, nw = X.size, Wt.size
    dX = torch.full((nx + guard,), float("nan"), device="cuda") if want_dx else None
    dW = torch.full((nw + guard,), float("nan"), device="cuda")
    ls.sconv2d_backward(dX, dW, torch.from_numpy(dY).cuda(), torch.from_numpy(Wt).cuda(),
                        torch.from_numpy(X).cuda(), batch, C, H, W, Cout, R, S, st, pad, dil, prec)
    torch.cuda.synchronize()
    gx = (dX[:nx].reshape(X.shape).cpu().numpy(), dX[nx:].cpu().numpy()) if want_dx else (None, None)
    return gx, (dW[:nw].reshape(Wt.shape).cpu().numpy(), dW[nw:].cpu().numpy())


def _bwd_case(g, prec, want_dx):
    batch, C, H, W, Cout, R, S, st, pad, dil = g
    X, Wt = _inputs(batch, C, H, W, Cout, R, S, seed=sum(g[:7]) + 1)
    P, Q = ls.conv2d_output_size(H, W, R, S, st, pad, dil)
    dY = np.random.default_rng(sum(g[:7])).uniform(-1, 1, (batch, Cout, P, Q)).astype(np.float32)
    (gx, tx), (gw, tw) = _run_bwd(X, Wt, dY, st, pad, dil, prec, want_dx)
    rx, ax, rw, aw = oracle.sconv2d_bwd_ref(X, Wt, dY, st, pad, dil)
    _assert_within(gw, rw, aw, batch * P * Q, prec, f"dW {g}")
    assert np.isnan(tw).all(), "write past the end of dW"
    if want_dx:
        _assert_within(gx, rx, ax, Cout * R * S, prec, f"dX {g}")
        assert np.isnan(tx).all(), "write past the end of dX"


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("g", BWD_GEOMS)
def test_conv_backward_matches_oracle(g, prec):
    print(f"conv backward {g} precision {prec}", flush=True)
    _bwd_case(g, prec, True)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("g", BWD_STRIDED)
def test_conv_backward_strided(g, prec):
    print(f"conv backward (strided) {g} precision {prec}", flush=True)
    _bwd_case(g, prec, True)


def _torch_grads(X, Wt, dY):
    x = torch.from_numpy(X).double().requires_grad_()
    w = torch.from_numpy(Wt).double().requires_grad_()
    torch.nn.functional.conv2d(x, w, padding=1).backward(torch.from_numpy(dY).double())
    xa = torch.from_numpy(X).double().abs().requires_grad_()
    wa = torch.from_numpy(Wt).double().abs().requires_grad_()
    torch.nn.functional.conv2d(xa, wa, padding=1).backward(torch.from_numpy(dY).double().abs())
    return x.grad.numpy(), xa.grad.numpy(), w.grad.numpy(), wa.grad.numpy()


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
@pytest.mark.parametrize("layer", [(64, 64, 56, 2), (128, 128, 28, 2), (256, 256, 14, 2), (512, 512, 7, 2),
                                   (256, 16, 32, 16), (512, 512, 7, 32)])
def test_conv_backward_layers_match_fp64(layer, prec):
    """The 3 x 3 layers against torch's float64 autograd (the |.| products
    for the bound are the gradients of the |.| problem).  The 7 x 7 layer at
    batch 32 takes the plan whose slices are twice the base length (one round
    of 248 units of 128 x 128 on 256 compute units)."""
    Cout, C, HW, batch = layer
    print(f"conv backward layer {layer} precision {prec}", flush=True)
    X, Wt = _inputs(batch, C, HW, HW, Cout, 3, 3, seed=Cout + C + 1)
    dY = np.random.default_rng(Cout).uniform(-1, 1, (batch, Cout, HW, HW)).astype(np.float32)
    (gx, _), (gw, _) = _run_bwd(X, Wt, dY, (1, 1), (1, 1), (1, 1), prec)
    rx, ax, rw, aw = _torch_grads(X, Wt, dY)
    _assert_within(gx, rx, ax, Cout * 9, prec, f"dX {layer}")
    _assert_within(gw, rw, aw, batch * HW * HW, prec, f"dW {layer}")


def test_conv_backward_batch_folded_few_channels():
    """The input gradient of small images folds the batch into one GEMM
    whose 16-row sub-tiles hold 2 live rows: the other rows' stores must be
    dropped (they would land on other images' pixels).  Repeated calls, each
    checked against the oracle and all bit-identical."""
    g = (24, 2, 5, 5, 3, 3, 3, (1, 1), (1, 1), (1, 1))
    batch, C, H, W, Cout, R, S, st, pad, dil = g
    X, Wt = _inputs(batch, C, H, W, Cout, R, S, seed=41)
    dY = np.random.default_rng(42).uniform(-1, 1, (batch, Cout, H, W)).astype(np.float32)
    rx, ax, _, _ = oracle.sconv2d_bwd_ref(X, Wt, dY, st, pad, dil)
    first = None
    for rep in range(8):
        (gx, tx), _ = _run_bwd(X, Wt, dY, st, pad, dil, ls.FP32)
        _assert_within(gx, rx, ax, Cout * R * S, ls.FP32, f"dX {g} call {rep}")
        assert np.isnan(tx).all(), "write past the end of dX"
        if first is None:
            first = gx
        assert np.array_equal(gx, first), f"call {rep} differs from the first"


def test_conv_backward_deterministic_and_single_gradients():
    """Sliced reductions summed in a fixed order: the same bits on every
    call, and each gradient alone equals the fused launch's."""
    X, Wt = _inputs(4, 128, 14, 14, 64, 3, 3, seed=21)
    dY = np.random.default_rng(5).uniform(-1, 1, (4, 64, 14, 14)).astype(np.float32)
    (x1, _), (w1, _) = _run_bwd(X, Wt, dY, (1, 1), (1, 1), (1, 1), ls.BF16X3)
    (x2, _), (w2, _) = _run_bwd(X, Wt, dY, (1, 1), (1, 1), (1, 1), ls.BF16X3)
    assert np.array_equal(x1, x2) and np.array_equal(w1, w2)
    (_, _), (w3, _) = _run_bwd(X, Wt, dY, (1, 1), (1, 1), (1, 1), ls.BF16X3, want_dx=False)
    assert np.isfinite(w3).all()
    dX = torch.full(X.shape, float("nan"), device="cuda")
    ls.sconv2d_backward(dX, None, torch.from_numpy(dY).cuda(), torch.from_numpy(Wt).cuda(), None, 4, 128, 14, 14,
                        64, 3, 3, (1, 1), (1, 1), (1, 1), ls.BF16X3)
    torch.cuda.synchronize()
    assert np.isfinite(dX.cpu().numpy()).all()
