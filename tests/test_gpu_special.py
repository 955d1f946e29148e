"""IEEE special values through every split-precision path.

The bf16 pieces of a special value are not IEEE-faithful on their own: the
later pieces of an Inf are NaN (Inf - Inf), a finite value above the
largest bf16 (about 3.3895e38) rounds its first piece to Inf, and a cross
product Inf x (a zero later piece of the other operand) is NaN where FP32
gives Inf.  The kernels recompute in FP32 every element whose split result
is Inf or NaN.  For every mode and entry point, with +-Inf, NaN, +-FLT_MAX,
values beyond the largest bf16 and subnormals planted in the operands:

* every output element has the class (finite, +Inf, -Inf or NaN) of the
  FP32 mode's output, and the FP32 mode's classes are those of the
  long-double oracle (the GEMMs; the convolutions' padding taps multiply
  stored zeros, as FP32 does, so only the classes across modes are compared
  there);
* every finite element is within the mode's bound of the oracle
  (|c - c_ref| <= (2 K u + extra) (|A||B|)_ij).

The planted values are placed so that no output sums two products that
overflow with opposite signs or lands within rounding of FLT_MAX (the
classes are then fixed by the mathematics, whatever the summation order):
A's large values sit in odd k, B's in even k, and the rows of A holding a
large value are zero where B holds one.
"""
import numpy as np
import pytest
import torch

import libshalom_amd as ls
import oracle

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
EXTRA = {ls.FP32: 0.0, ls.BF16X6: 0.0, ls.BF16X3: 3 * 2.0 ** -16, ls.BF16: 2 * 2.0 ** -8}
SPLIT = [ls.BF16X3, ls.BF16X6, ls.BF16]
FMAX = float(np.finfo(np.float32).max)
INF = float("inf")
NAN = float("nan")

# (row of op(A), k, value); k odd
A_SPECIAL = [(0, 1, INF), (2, 3, -INF), (4, 5, NAN), (6, 1, FMAX), (8, 3, -FMAX), (10, 5, 3.4e38),
             (12, 7, -3.395e38), (14, 1, 1e-40), (16, 3, 1.4e-45)]
# (k, column of op(B), value); k even, except the zeros that meet A's Infs
B_SPECIAL = [(1, 2, 0.0), (3, 3, 0.0), (2, 4, INF), (4, 5, NAN), (6, 6, 3.39e38), (8, 7, -3.4e38),
             (10, 8, 1e-39), (12, 9, -2e-44)]


def classes(x):
    """0 finite, 1 +Inf, 2 -Inf, 3 NaN (of the float32 value)."""
    x = np.asarray(x, dtype=np.float64).astype(np.float32)
    c = np.zeros(x.shape, dtype=np.int8)
    c[np.isposinf(x)] = 1
    c[np.isneginf(x)] = 2
    c[np.isnan(x)] = 3
    return c


def planted(M, N, K, seed):
    """op(A) (M x K) and op(B) (K x N) with the special values planted."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    big_b_k = [k for k, j, v in B_SPECIAL if k < K and j < N and abs(v) > 1e30]
    for i, k, v in A_SPECIAL:
        if i < M and k < K:
            A[i, k] = v
            if abs(v) > 1e30 and np.isfinite(v):
                A[i, big_b_k] = 0.0
    for k, j, v in B_SPECIAL:
        if k < K and j < N:
            B[k, j] = v
    return A, B


def stored(X, trans):
    return np.ascontiguousarray(X.T if trans else X)


def check_gemm(C, A, B, ta, tb, M, N, K, prec, C32=None, alpha=1.0, beta=0.0, C0=None):
    """Classes against the oracle (and FP32's output), finite values against the bound."""
    Cref, absAB = oracle.sgemm_ref(ta, tb, A, B, M, N, K)
    if alpha != 1.0 or beta != 0.0:
        Cref = alpha * Cref + beta * C0.astype(np.float64)
        absAB = abs(alpha) * absAB + abs(beta) * np.abs(C0.astype(np.float64))
    cls = classes(C)
    if C32 is not None:
        assert np.array_equal(cls, classes(C32)), np.argwhere(cls != classes(C32))[:8]
    assert np.array_equal(cls, classes(Cref)), np.argwhere(cls != classes(Cref))[:8]
    fin = (cls == 0) & np.isfinite(Cref) & np.isfinite(absAB)
    bound = (2 * max(K, 1) * U32 + EXTRA[prec]) * absAB
    err = np.abs(C.astype(np.float64) - Cref)
    assert (err[fin] <= bound[fin]).all(), np.argwhere(fin & (err > bound))[:8]
    # the planted specials reach the outputs they should
    assert (cls != 0).sum() > M + N


def run_ex(dev, ta, tb, M, N, K, prec, mp):
    A, B = planted(M, N, K, seed=M * 7 + N * 3 + K)
    dA = torch.from_numpy(stored(A, ta)).to(dev)
    dB = torch.from_numpy(stored(B, tb)).to(dev)
    dC = torch.full((M, N), NAN, device=dev)
    (ls.sgemm_mp_ex if mp else ls.sgemm_ex)(ta, tb, dC, dA, dB, M, N, K, prec)
    return dC.cpu().numpy(), stored(A, ta), stored(B, tb)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_sgemm_special_values(dev, ta, tb):
    """One tile per workgroup (96 x 80 x 64: 32 x 32 tiles)."""
    M, N, K = 96, 80, 64
    C32, A, B = run_ex(dev, ta, tb, M, N, K, ls.FP32, mp=False)
    check_gemm(C32, A, B, ta, tb, M, N, K, ls.FP32)
    for prec in SPLIT:
        C, _, _ = run_ex(dev, ta, tb, M, N, K, prec, mp=False)
        check_gemm(C, A, B, ta, tb, M, N, K, prec, C32)


@pytest.mark.parametrize("shape", [(2048, 16, 2048), (512, 384, 160)])
def test_sgemm_mp_special_values(dev, shape):
    """Split-K (the partial tiles are repaired before they are summed) and 128-row tiles."""
    M, N, K = shape
    C32, A, B = run_ex(dev, 0, 0, M, N, K, ls.FP32, mp=True)
    check_gemm(C32, A, B, 0, 0, M, N, K, ls.FP32)
    for prec in SPLIT:
        C, _, _ = run_ex(dev, 0, 0, M, N, K, prec, mp=True)
        check_gemm(C, A, B, 0, 0, M, N, K, prec, C32)


def test_sgemm_general_special_values_alpha_beta(dev):
    """alpha / beta applied once to the recomputed elements, C read for beta."""
    M, N, K = 96, 80, 64
    A, B = planted(M, N, K, seed=11)
    rng = np.random.default_rng(12)
    C0 = rng.uniform(-1, 1, (M, N)).astype(np.float32)
    out = {}
    for prec in [ls.FP32] + SPLIT:
        dC = torch.from_numpy(C0.copy()).to(dev)
        ls.sgemm_general(0, 0, M, N, K, 1.5, torch.from_numpy(A).to(dev), K, torch.from_numpy(B).to(dev), N,
                         0.5, dC, N, prec)
        out[prec] = dC.cpu().numpy()
        check_gemm(out[prec], A, B, 0, 0, M, N, K, prec, out.get(ls.FP32), alpha=1.5, beta=0.5, C0=C0)


def test_batched_special_values(dev):
    """Persistent batched kernels: the jobs of a stream that held one are read back after it."""
    shapes = [(96, 80, 64), (33, 17, 21), (128, 128, 128), (200, 40, 96), (20, 20, 20), (130, 70, 33)]
    As, Bs = [], []
    for n, (M, N, K) in enumerate(shapes):
        if n % 2 == 0:
            A, B = planted(M, N, K, seed=100 + n)
        else:  # no special value: these tiles must stay as computed
            rng = np.random.default_rng(100 + n)
            A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
            B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        As.append(A)
        Bs.append(B)
    dA = [torch.from_numpy(a).to(dev) for a in As]
    dB = [torch.from_numpy(b).to(dev) for b in Bs]
    Ms, Ns, Ks = zip(*shapes)
    res = {}
    for prec in [ls.FP32] + SPLIT:
        dC = [torch.full((M, N), NAN, device=dev) for M, N, _ in shapes]
        ls.sgemm_batched(0, 0, dC, dA, dB, Ms, Ns, Ks, prec)
        res[prec] = [c.cpu().numpy() for c in dC]
    for n, (M, N, K) in enumerate(shapes):
        for prec in [ls.FP32] + SPLIT:
            C = res[prec][n]
            if n % 2 == 0:
                check_gemm(C, As[n], Bs[n], 0, 0, M, N, K, prec, res[ls.FP32][n])
            else:
                Cref, absAB = oracle.sgemm_ref(0, 0, As[n], Bs[n], M, N, K)
                bound = (2 * K * U32 + EXTRA[prec]) * absAB
                assert (np.abs(C - Cref) <= bound).all()


def test_strided_and_backward_special_values(dev):
    """The slice engine: strided batch forward and the fused backward (dA, dB), plain operands."""
    M, N, K, batch = 64, 48, 80, 3
    As, Bs = zip(*[planted(M, N, K, seed=200 + b) for b in range(batch)])
    A = np.stack(As)
    B = np.stack(Bs)
    # backward of one GEMM with the specials in the upstream gradient G (M x N)
    G = planted(M, 1, N, seed=300)[0]
    rng = np.random.default_rng(301)
    Ab = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    Bb = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    dev_ = [torch.from_numpy(x).to(dev) for x in (A, B, G, Ab, Bb)]
    fwd, bwd = {}, {}
    for prec in [ls.FP32] + SPLIT:
        dC = torch.full((batch, M, N), NAN, device=dev)
        ls.sgemm_strided_batched(0, 0, dC, dev_[0], dev_[1], M, N, K, batch, M * K, K * N, M * N, prec)
        fwd[prec] = dC.cpu().numpy()
        gA = torch.full((M, K), NAN, device=dev)
        gB = torch.full((K, N), NAN, device=dev)
        ls.sgemm_backward(0, 0, dev_[2], dev_[3], dev_[4], gA, gB, M, N, K, prec)
        bwd[prec] = (gA.cpu().numpy(), gB.cpu().numpy())
    for prec in [ls.FP32] + SPLIT:
        for b in range(batch):
            check_gemm(fwd[prec][b], A[b], B[b], 0, 0, M, N, K, prec, fwd[ls.FP32][b])
        gA, gB = bwd[prec]
        # dA = G (M x N) B^T (N x K); dB = A^T (K x M) G (M x N)
        check_gemm(gA, G, Bb, 0, 1, M, K, N, prec, bwd[ls.FP32][0])
        check_gemm(gB, Ab, G, 1, 0, K, N, M, prec, bwd[ls.FP32][1])


def _conv_special_inputs(batch, C, H, W, Cout, R, S, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (batch, C, H, W)).astype(np.float32)
    Wt = rng.uniform(-1, 1, (Cout, C, R, S)).astype(np.float32)
    X[0, 0, 2 % H, 3 % W] = INF
    X[0, 1, 5 % H, 1 % W] = NAN
    X[-1, 0, 4 % H, 4 % W] = 3.4e38
    X[-1, 1, 1 % H, 6 % W] = -FMAX   # never in one receptive field with the one above
    X[0, 0, 6 % H, 6 % W] = 1e-40
    return X, Wt


@pytest.mark.parametrize("geom", [
    (2, 3, 9, 11, 5, 3, 3, (1, 1), (1, 1), (1, 1)),
    (2, 4, 12, 12, 70, 3, 3, (2, 2), (1, 1), (1, 1)),    # strided: dX fractionally strided
    (2, 3, 10, 12, 4, 3, 7, (1, 1), (1, 3), (1, 1)),     # dW's gather in 16-byte runs
    (8, 16, 7, 7, 32, 3, 3, (1, 1), (1, 1), (1, 1)),     # 49 pixels an image: the batch folded
])
def test_conv_special_values(dev, geom):
    """Implicit-GEMM convolution forward and backward: the recompute gathers the im2col
    elements as the loads do (padding, strides, dilation, the folded batch)."""
    batch, C, H, W, Cout, R, S, st, pad, dil = geom
    X, Wt = _conv_special_inputs(batch, C, H, W, Cout, R, S, seed=sum(geom[:7]))
    P, Q = ls.conv2d_output_size(H, W, R, S, st, pad, dil)
    rng = np.random.default_rng(5)
    dY = rng.uniform(-1, 1, (batch, Cout, P, Q)).astype(np.float32)
    dY[0, 0, 1, 1] = INF
    dY[-1, Cout - 1, P - 1, 0] = NAN
    dY[0, Cout // 2, 0, Q - 1] = 3.4e38
    Xd, Wd, dYd = (torch.from_numpy(a).to(dev) for a in (X, Wt, dY))
    Yref, aY = oracle.sconv2d_ref(X, Wt, st, pad, dil)
    dXref, adX, dWref, adW = oracle.sconv2d_bwd_ref(X, Wt, dY, st, pad, dil)
    out = {}
    for prec in [ls.FP32] + SPLIT:
        Y = torch.full((batch, Cout, P, Q), NAN, device=dev)
        ls.sconv2d_forward(Y, Wd, Xd, batch, C, H, W, Cout, R, S, st, pad, dil, prec)
        dX = torch.full(X.shape, NAN, device=dev)
        dW = torch.full(Wt.shape, NAN, device=dev)
        ls.sconv2d_backward(dX, dW, dYd, Wd, Xd, batch, C, H, W, Cout, R, S, st, pad, dil, prec)
        out[prec] = [t.cpu().numpy() for t in (Y, dX, dW)]
    # reduction lengths: forward C*R*S; dX at most Cout*R*S; dW batch*P*Q
    Ks = [C * R * S, Cout * R * S, batch * P * Q]
    for prec in [ls.FP32] + SPLIT:
        for got, want, ref, absp, K in zip(out[prec], out[ls.FP32], (Yref, dXref, dWref), (aY, adX, adW), Ks):
            cg = classes(got)
            assert np.array_equal(cg, classes(want)), np.argwhere(cg != classes(want))[:8]
            assert (cg != 0).any()
            fin = (cg == 0) & np.isfinite(ref) & np.isfinite(absp)
            bound = (2 * K * U32 + EXTRA[prec]) * absp
            assert (np.abs(got - ref)[fin] <= bound[fin]).all(), np.argwhere(fin & (np.abs(got - ref) > bound))[:8]
