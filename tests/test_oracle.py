"""Pins the CPU oracle before anything is checked against it (no GPU).

* drand48 / input recipe: against glibc's own drand48 (golden fixture written
  by tests/golden/make_golden.py through ctypes, not through the oracle).
* Reference GEMM: against numpy float64 matmul on the golden inputs, exact
  integer products (all trans combinations, non-square), the all-ones
  known answer of the reference harnesses (evaluate/test.c:38-46), and
  transpose identities (NT(A,B) == NN(A,B^T)).
* Checker: planted mistakes (transposed operand, one perturbed element, NaN,
  dropped k term) must be caught; the exactly rounded result must pass.
* Reference convolution: against torch's float64 conv2d (stride, padding,
  dilation, non-square filters), against torch's unfold (im2col) followed by
  the reference GEMM (bit-identical), and by closed forms: a one-tap filter
  copies or shifts the input, all-ones operands count the taps inside the
  image.  A filter flipped along one axis must be caught.
* Convolution gradients: against torch's float64 autograd, and the adjoint
  identities <conv(X, W), dY> = <X, dX> = <W, dW>.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _unhex(xs):
    return np.array([float.fromhex(x) for x in xs], dtype=np.float64)


def test_drand48_matches_glibc():
    g = _load("drand48_first.json")
    r = oracle.Drand48()
    got = r.doubles(len(g["drand48"]))
    assert np.array_equal(got, _unhex(g["drand48"]))


def test_recipe_matches_glibc_recipe():
    g = _load("drand48_first.json")
    r = oracle.Drand48()
    got = r.matrix(4, 4).ravel()
    assert np.array_equal(got.astype(np.float64), _unhex(g["recipe_f32"]))
    assert got.min() >= -1.0 and got.max() < 1.0


def test_hello_fixture():
    g = _load("hello_80_nt.json")
    M, N, K = g["M"], g["N"], g["K"]
    A, B = oracle.hello_inputs(M, N, K)
    Cref, absAB = oracle.sgemm_ref(0, 1, A, B, M, N, K)
    np.testing.assert_allclose(Cref[:4, :4].ravel(), _unhex(g["corner"]), rtol=0, atol=1e-13)
    np.testing.assert_allclose(np.diag(Cref), _unhex(g["diag"]), rtol=0, atol=1e-13)
    assert abs(Cref.sum() - float.fromhex(g["sum"])) < 1e-10
    assert abs(np.abs(Cref).sum() - float.fromhex(g["abs_sum"])) < 1e-10


def test_tiny_cases_fixture():
    cases = _load("tiny_cases.json")
    r = oracle.Drand48()
    for c in cases:
        M, N, K, ta, tb = c["M"], c["N"], c["K"], c["transa"], c["transb"]
        A = r.matrix(1, M * K).ravel()
        B = r.matrix(1, K * N).ravel()
        assert np.array_equal(A.astype(np.float64), _unhex(c["A"]))
        assert np.array_equal(B.astype(np.float64), _unhex(c["B"]))
        Cref, absAB = oracle.sgemm_ref(ta, tb, A, B, M, N, K)
        np.testing.assert_allclose(Cref.ravel(), _unhex(c["C"]), rtol=0, atol=1e-14 * max(1, K))


@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
def test_integer_products_exact(ta, tb):
    rng = np.random.default_rng(7 + 2 * ta + tb)
    M, N, K = 7, 11, 5
    A = rng.integers(-8, 9, size=oracle.a_shape(ta, M, K)).astype(np.float32)
    B = rng.integers(-8, 9, size=oracle.b_shape(tb, N, K)).astype(np.float32)
    opA = A.T if ta else A
    opB = B.T if tb else B
    exact = opA.astype(np.int64) @ opB.astype(np.int64)
    Cref, absAB = oracle.sgemm_ref(ta, tb, A, B, M, N, K)
    assert np.array_equal(Cref, exact.astype(np.float64))
    assert np.array_equal(absAB, (np.abs(opA).astype(np.int64) @ np.abs(opB).astype(np.int64)).astype(np.float64))
    Cd, _ = oracle.dgemm_ref(ta, tb, A.astype(np.float64), B.astype(np.float64), M, N, K)
    assert np.array_equal(Cd, exact.astype(np.float64))


@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
def test_all_ones_known_answer(ta, tb):
    M, N, K = 9, 13, 17
    A = np.ones(M * K, dtype=np.float32)
    B = np.ones(K * N, dtype=np.float32)
    Cref, absAB = oracle.sgemm_ref(ta, tb, A, B, M, N, K)
    assert (Cref == K).all() and (absAB == K).all()


def test_transpose_identity():
    r = oracle.Drand48(12345)
    M, N, K = 6, 10, 8
    A = r.matrix(M, K)
    B = r.matrix(K, N)
    c_nn, _ = oracle.sgemm_ref(0, 0, A, B, M, N, K)
    c_nt, _ = oracle.sgemm_ref(0, 1, A, np.ascontiguousarray(B.T), M, N, K)
    c_tn, _ = oracle.sgemm_ref(1, 0, np.ascontiguousarray(A.T), B, M, N, K)
    c_tt, _ = oracle.sgemm_ref(1, 1, np.ascontiguousarray(A.T), np.ascontiguousarray(B.T), M, N, K)
    assert np.array_equal(c_nn, c_nt) and np.array_equal(c_nn, c_tn) and np.array_equal(c_nn, c_tt)
    np.testing.assert_allclose(c_nn, A.astype(np.float64) @ B.astype(np.float64), rtol=0, atol=1e-13)


def test_checker_catches_planted_mistakes():
    r = oracle.Drand48(99)
    M, N, K = 20, 24, 40
    A = r.matrix(M, K)
    B = r.matrix(K, N)
    Cref, absAB = oracle.sgemm_ref(0, 0, A, B, M, N, K)
    good = Cref.astype(np.float32)
    bad, st = oracle.check(good, Cref, absAB, K)
    assert bad == 0 and st["max_ratio"] <= 1.0
    # FP32 triple loop is within tolerance
    cpu = oracle.cpu_sgemm(0, 0, A, B, M, N, K)
    assert oracle.check(cpu, Cref, absAB, K)[0] == 0
    # transposed operand (square sub-problem so shapes line up)
    At = np.ascontiguousarray(A[:, :20].T)
    Cw, _ = oracle.sgemm_ref(0, 0, At, B[:20], 20, N, 20)
    Cr2, S2 = oracle.sgemm_ref(0, 0, A[:, :20].copy(), B[:20].copy(), 20, N, 20)
    assert oracle.check(Cw.astype(np.float32), Cr2, S2, 20)[0] > 0
    # dropped last k term
    Cd, _ = oracle.sgemm_ref(0, 0, A[:, :K - 1].copy(), B[:K - 1].copy(), M, N, K - 1)
    assert oracle.check(Cd.astype(np.float32), Cref, absAB, K)[0] > 0
    # one element off by 10x the bound
    p = good.copy()
    p[3, 5] += np.float32(20 * K * 2.0 ** -24 * absAB[3, 5])
    assert oracle.check(p, Cref, absAB, K)[0] == 1
    # NaN
    p = good.copy()
    p[0, 0] = np.nan
    assert oracle.check(p, Cref, absAB, K)[0] == 1


def test_flat_reference_and_kv_check():
    r = oracle.Drand48(5)
    shapes = [(3, 4, 5), (1, 1, 1), (8, 2, 3)]
    As = [r.matrix(m, k).ravel() for m, n, k in shapes]
    Bs = [r.matrix(k, n).ravel() for m, n, k in shapes]
    offA = np.cumsum([0] + [a.size for a in As])[:-1]
    offB = np.cumsum([0] + [b.size for b in Bs])[:-1]
    offC = np.cumsum([0] + [m * n for m, n, k in shapes])
    Cref, S, Ke = oracle.gemm_ref_flat(0, 0, np.concatenate(As), np.concatenate(Bs), offA, offB, offC[:-1],
                                       [s[0] for s in shapes], [s[1] for s in shapes], [s[2] for s in shapes],
                                       int(offC[-1]))
    for g, (m, n, k) in enumerate(shapes):
        c1, s1 = oracle.sgemm_ref(0, 0, As[g], Bs[g], m, n, k)
        assert np.array_equal(Cref[offC[g]:offC[g + 1]], c1.ravel())
        assert (Ke[offC[g]:offC[g + 1]] == k).all()
    bad, _ = oracle.check_kv(Cref.astype(np.float32), Cref, S, Ke)
    assert bad == 0
    wrong = Cref.astype(np.float32).copy()
    wrong[-1] += 1.0
    assert oracle.check_kv(wrong, Cref, S, Ke)[0] == 1


def test_cpu_baseline_batched_matches_single():
    r = oracle.Drand48(77)
    shapes = [(5, 7, 3), (16, 1, 9), (2, 2, 2)]
    for ta in (0, 1):
        for tb in (0, 1):
            As = [r.matrix(*oracle.a_shape(ta, m, k)) for m, n, k in shapes]
            Bs = [r.matrix(*oracle.b_shape(tb, n, k)) for m, n, k in shapes]
            Cs = [np.zeros((m, n), np.float32) for m, n, k in shapes]
            oracle.cpu_gemm_batched(ta, tb, Cs, As, Bs, *zip(*shapes))
            for g, (m, n, k) in enumerate(shapes):
                single = oracle.cpu_sgemm(ta, tb, As[g], Bs[g], m, n, k)
                assert np.array_equal(Cs[g], single)
                cref, s = oracle.sgemm_ref(ta, tb, As[g], Bs[g], m, n, k)
                assert oracle.check(single, cref, s, k)[0] == 0


def test_workload_hello_inputs_match_oracle():
    """The package's drand48 input generator (used by bench.py) reproduces the
    oracle's Hello.c inputs bit for bit."""
    from libshalom_amd.workloads import hello_inputs
    A, B = hello_inputs(80, 80, 80)
    A0, B0 = oracle.hello_inputs(80, 80, 80)
    assert np.array_equal(A, A0) and np.array_equal(B, B0)


# (batch, C, H, W, Cout, R, S, stride, padding, dilation)
CONV_GEOMS = [
    (2, 3, 9, 11, 5, 3, 3, (1, 1), (1, 1), (1, 1)),
    (1, 4, 8, 8, 6, 1, 1, (1, 1), (0, 0), (1, 1)),
    (2, 5, 13, 7, 6, 3, 2, (2, 1), (1, 0), (1, 1)),
    (1, 3, 15, 15, 4, 3, 3, (1, 1), (2, 2), (2, 2)),
    (3, 2, 6, 5, 3, 5, 5, (1, 1), (3, 3), (1, 1)),
    (1, 2, 5, 9, 2, 2, 4, (3, 2), (0, 1), (1, 2)),
]


def _conv_inputs(g, seed):
    batch, C, H, W, Cout, R, S = g[:7]
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (batch, C, H, W)).astype(np.float32)
    Wt = rng.uniform(-1, 1, (Cout, C, R, S)).astype(np.float32)
    return X, Wt


@pytest.mark.parametrize("g", CONV_GEOMS)
def test_conv_reference_matches_torch(g):
    torch = pytest.importorskip("torch")
    X, Wt = _conv_inputs(g, 7)
    st, pad, dil = g[7:]
    Y, A = oracle.sconv2d_ref(X, Wt, st, pad, dil)
    Yt = torch.nn.functional.conv2d(torch.from_numpy(X).double(), torch.from_numpy(Wt).double(), stride=st,
                                    padding=pad, dilation=dil).numpy()
    At = torch.nn.functional.conv2d(torch.from_numpy(X).double().abs(), torch.from_numpy(Wt).double().abs(),
                                    stride=st, padding=pad, dilation=dil).numpy()
    assert Y.shape == Yt.shape
    np.testing.assert_allclose(Y, Yt, rtol=0, atol=1e-13)
    np.testing.assert_allclose(A, At, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("g", CONV_GEOMS)
def test_conv_reference_is_gemm_of_unfold(g):
    torch = pytest.importorskip("torch")
    X, Wt = _conv_inputs(g, 11)
    batch, C, H, W, Cout, R, S = g[:7]
    st, pad, dil = g[7:]
    Y, _ = oracle.sconv2d_ref(X, Wt, st, pad, dil)
    cols = torch.nn.functional.unfold(torch.from_numpy(X), (R, S), dilation=dil, padding=pad, stride=st).numpy()
    K, PQ = C * R * S, Y.shape[2] * Y.shape[3]
    assert cols.shape == (batch, K, PQ)
    for b in range(batch):
        Cb, _ = oracle.sgemm_ref(0, 0, Wt.reshape(Cout, K), np.ascontiguousarray(cols[b]), Cout, PQ, K)
        assert np.array_equal(Cb, Y[b].reshape(Cout, PQ))


def test_conv_reference_one_tap_filters():
    X, _ = _conv_inputs((2, 3, 7, 9, 3, 3, 3), 3)
    Wt = np.zeros((3, 3, 3, 3), dtype=np.float32)
    for o in range(3):
        Wt[o, o, 1, 1] = 1.0  # centre tap, padding 1: the identity
    Y, _ = oracle.sconv2d_ref(X, Wt, (1, 1), (1, 1), (1, 1))
    assert np.array_equal(Y, X.astype(np.float64))
    Wt[:] = 0
    for o in range(3):
        Wt[o, (o + 1) % 3, 0, 0] = 1.0  # top-left tap: shift down-right by one, channel o+1
    Y, _ = oracle.sconv2d_ref(X, Wt, (1, 1), (1, 1), (1, 1))
    for o in range(3):
        src = X[:, (o + 1) % 3].astype(np.float64)
        assert np.array_equal(Y[:, o, 1:, 1:], src[:, :-1, :-1])
        assert not Y[:, o, 0, :].any() and not Y[:, o, :, 0].any()
    # stride 2, no padding: every other pixel
    Wt[:] = 0
    Wt[0, 0, 0, 0] = 1.0
    Y, _ = oracle.sconv2d_ref(X, Wt[:1, :, :1, :1].copy(), (2, 2), (0, 0), (1, 1))
    assert np.array_equal(Y[:, 0], X[:, 0, ::2, ::2].astype(np.float64))


def test_conv_reference_all_ones_counts_taps():
    C, H, W = 4, 6, 5
    X = np.ones((1, C, H, W), dtype=np.float32)
    Wt = np.ones((2, C, 3, 3), dtype=np.float32)
    Y, A = oracle.sconv2d_ref(X, Wt, (1, 1), (1, 1), (1, 1))
    rows = np.array([2] + [3] * (H - 2) + [2], dtype=np.float64)
    cols = np.array([2] + [3] * (W - 2) + [2], dtype=np.float64)
    expect = C * np.outer(rows, cols)
    assert np.array_equal(Y[0, 0], expect) and np.array_equal(Y[0, 1], expect) and np.array_equal(A[0, 0], expect)


def test_conv_reference_catches_flipped_filter():
    g = CONV_GEOMS[0]
    X, Wt = _conv_inputs(g, 5)
    Y, A = oracle.sconv2d_ref(X, Wt, (1, 1), (1, 1), (1, 1))
    Yf, _ = oracle.sconv2d_ref(X, np.ascontiguousarray(Wt[:, :, :, ::-1]), (1, 1), (1, 1), (1, 1))
    K = g[1] * g[5] * g[6]
    assert oracle.check(Y.astype(np.float32), Y, A, K)[0] == 0
    assert oracle.check(Yf.astype(np.float32), Y, A, K)[0] > 0


@pytest.mark.parametrize("g", CONV_GEOMS)
def test_conv_backward_reference_matches_torch_autograd(g):
    torch = pytest.importorskip("torch")
    X, Wt = _conv_inputs(g, 13)
    st, pad, dil = g[7:]
    Y, _ = oracle.sconv2d_ref(X, Wt, st, pad, dil)
    dY = np.random.default_rng(3).uniform(-1, 1, Y.shape).astype(np.float32)
    dX, aX, dW, aW = oracle.sconv2d_bwd_ref(X, Wt, dY, st, pad, dil)
    x = torch.from_numpy(X).double().requires_grad_()
    w = torch.from_numpy(Wt).double().requires_grad_()
    y = torch.nn.functional.conv2d(x, w, stride=st, padding=pad, dilation=dil)
    y.backward(torch.from_numpy(dY).double())
    np.testing.assert_allclose(dX, x.grad.numpy(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(dW, w.grad.numpy(), rtol=0, atol=1e-12)
    assert (aX >= np.abs(dX) - 1e-12).all() and (aW >= np.abs(dW) - 1e-12).all()


@pytest.mark.parametrize("g", CONV_GEOMS[:3])
def test_conv_backward_reference_adjoint_identities(g):
    X, Wt = _conv_inputs(g, 17)
    st, pad, dil = g[7:]
    Y, _ = oracle.sconv2d_ref(X, Wt, st, pad, dil)
    dY = np.random.default_rng(4).uniform(-1, 1, Y.shape).astype(np.float32)
    dX, _, dW, _ = oracle.sconv2d_bwd_ref(X, Wt, dY, st, pad, dil)
    lhs = float(np.sum(Y * dY.astype(np.float64)))
    assert abs(lhs - float(np.sum(X.astype(np.float64) * dX))) < 1e-10
    assert abs(lhs - float(np.sum(Wt.astype(np.float64) * dW))) < 1e-10
