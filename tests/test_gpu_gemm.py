"""GPU parity tests: every entry point through the C-ABI against the oracle.

Bar: |C - C_ref| <= 2*K*u*(|A||B|)_ij elementwise (u = 2^-24 for FP32 output,
2^-53 for FP64), C prefilled with NaN so every element must be written.
FP32 and BF16X6 are held to that bound exactly; BF16X3 and BF16, whose
two- and one-piece splits cannot meet it, get their documented extra
per-product term (PREC_EXTRA).  Inputs follow the reference harness recipe 2*(float)drand48()-1
from seeded drand48 streams.
"""
import numpy as np
import pytest
import torch

import libshalom_amd as ls
import oracle

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
U64 = 2.0 ** -53
# Extra relative error per product |a||b| allowed beyond 2*K*u (bf16 keeps 8
# significand bits, so a two-piece split leaves ~2^-16 of each operand and a
# one-piece split ~2^-8).  The three-piece split of BF16X6 represents an FP32
# value exactly and drops only the products of its two smaller pieces (at
# most ~2^-24 |a||b| together), which the 2*K*u bound absorbs: it gets none.
PREC_EXTRA = {ls.FP32: 0.0, ls.BF16X6: 0.0, ls.BF16X3: 3 * 2.0 ** -16, ls.BF16: 2 * 2.0 ** -8}
TRANS = [(0, 0), (0, 1), (1, 0), (1, 1)]


def dev_tensor(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def nan_c(M, N, dev, dtype=torch.float32):
    return torch.full((M, N), float("nan"), dtype=dtype, device=dev)


def check_f32(C, A, B, ta, tb, M, N, K, prec=ls.FP32):
    Cref, absAB = oracle.sgemm_ref(ta, tb, A, B, M, N, K)
    Ke = np.full(M * N, float(max(K, 1)))
    bad, st = oracle.check_kv(C.reshape(-1), Cref.reshape(-1), absAB.reshape(-1), Ke, U32, PREC_EXTRA[prec])
    return bad, st


def test_hello_80_nt(dev):
    """Hello.c: LibShalom_sgemm(NoTrans, Trans, C, A, B, 80, 80, 80)."""
    M = N = K = 80
    A, B = oracle.hello_inputs(M, N, K)
    dC = nan_c(M, N, dev)
    ls.sgemm(ls.NoTrans, ls.Trans, dC, dev_tensor(A, dev), dev_tensor(B, dev), M, N, K)
    C = dC.cpu().numpy()
    bad, st = check_f32(C, A, B, 0, 1, M, N, K)
    assert bad == 0, st
    assert np.isfinite(C).all()


def test_reference_api_host_pointers(dev):
    """The reference's own ABI: Hello.c passes host buffers.  The reference-API
    calls stage host operands through the device (all host, and mixed)."""
    M = N = K = 80
    A, B = oracle.hello_inputs(M, N, K)
    C = np.full((M, N), np.nan, dtype=np.float32)
    ls.sgemm(ls.NoTrans, ls.Trans, C, A, B, M, N, K)
    bad, st = check_f32(C, A, B, 0, 1, M, N, K)
    assert bad == 0, st
    # A on the device, B and C on the host; then C on the device only
    C2 = np.full((M, N), np.nan, dtype=np.float32)
    ls.sgemm(ls.NoTrans, ls.Trans, C2, dev_tensor(A, dev), B, M, N, K)
    assert np.array_equal(C2, C)
    dC = nan_c(M, N, dev)
    ls.sgemm(ls.NoTrans, ls.Trans, dC, A, B, M, N, K)
    assert np.array_equal(dC.cpu().numpy(), C)
    # sgemm_mp, tall-skinny, host operands
    r = oracle.Drand48()
    M2, N2, K2 = 4096, 16, 64
    A2, B2 = r.matrix(M2, K2), r.matrix(K2, N2)
    C3 = np.full((M2, N2), np.nan, dtype=np.float32)
    ls.sgemm_mp(ls.NoTrans, ls.NoTrans, C3, A2, B2, M2, N2, K2)
    bad, st = check_f32(C3, A2, B2, 0, 0, M2, N2, K2)
    assert bad == 0, st
    # dgemm with host float64 operands (a CP2K block shape)
    M3, N3, K3 = 23, 26, 13
    A3 = r.matrix(M3, K3).astype(np.float64)
    B3 = r.matrix(K3, N3).astype(np.float64)
    C4 = np.full((M3, N3), np.nan)
    ls.dgemm(ls.NoTrans, ls.NoTrans, C4, A3, B3, M3, N3, K3)
    Cref, absAB = oracle.dgemm_ref(0, 0, A3, B3, M3, N3, K3)
    assert (np.abs(C4 - Cref.reshape(M3, N3)) <= 2 * K3 * U64 * absAB.reshape(M3, N3) + 1e-300).all()


@pytest.mark.parametrize("ta,tb", TRANS)
def test_small_sweep_8_to_128(dev, ta, tb):
    """benchmark/small_SGEMM: M = N = K = 8..128 step 8 (all four trans combos)."""
    r = oracle.Drand48()
    for n in range(8, 129, 8):
        A = r.matrix(*oracle.a_shape(ta, n, n))
        B = r.matrix(*oracle.b_shape(tb, n, n))
        dC = nan_c(n, n, dev)
        ls.sgemm(ta, tb, dC, dev_tensor(A, dev), dev_tensor(B, dev), n, n, n)
        bad, st = check_f32(dC.cpu().numpy(), A, B, ta, tb, n, n, n)
        assert bad == 0, (n, st)


def _flat_batch(shapes, ta, tb, seed, f64=False):
    r = oracle.Drand48(seed)
    asz = [m * k for m, n, k in shapes]
    bsz = [k * n for m, n, k in shapes]
    csz = [m * n for m, n, k in shapes]
    offA = np.concatenate([[0], np.cumsum(asz)]).astype(np.int64)
    offB = np.concatenate([[0], np.cumsum(bsz)]).astype(np.int64)
    offC = np.concatenate([[0], np.cumsum(csz)]).astype(np.int64)
    if f64:
        A = r.matrix_d(1, int(offA[-1])).ravel()
        B = r.matrix_d(1, int(offB[-1])).ravel()
    else:
        A = r.matrix(1, int(offA[-1])).ravel()
        B = r.matrix(1, int(offB[-1])).ravel()
    return A, B, offA, offB, offC


def _run_batched(dev, shapes, ta, tb, prec, seed, f64=False, replays=1):
    A, B, offA, offB, offC = _flat_batch(shapes, ta, tb, seed, f64)
    dt = torch.float64 if f64 else torch.float32
    dA = dev_tensor(A, dev)
    dB = dev_tensor(B, dev)
    dC = torch.full((int(offC[-1]) + 1,), float("nan"), dtype=dt, device=dev)
    es = dA.element_size()
    Cs = [dC.data_ptr() + es * int(o) for o in offC[:-1]]
    As = [dA.data_ptr() + es * int(o) for o in offA[:-1]]
    Bs = [dB.data_ptr() + es * int(o) for o in offB[:-1]]
    Ms, Ns, Ks = zip(*shapes)
    outs = []
    with ls.BatchedPlan(ta, tb, Cs, As, Bs, Ms, Ns, Ks, precision=prec, fp64=f64) as plan:
        for _ in range(replays):
            dC.fill_(float("nan"))
            plan.execute()
            torch.cuda.synchronize()
            outs.append(dC.cpu().numpy().copy())
    Cref, S, Ke = oracle.gemm_ref_flat(ta, tb, A, B, offA[:-1], offB[:-1], offC[:-1], Ms, Ns, Ks, int(offC[-1]))
    return outs, Cref, S, Ke, offC


@pytest.mark.parametrize("ta,tb", TRANS)
def test_tiny_shapes_brute_force(dev, ta, tb):
    """Every (M, N, K) in [1..33]^3 (one batched plan per trans combo)."""
    shapes = [(m, n, k) for m in range(1, 34) for n in range(1, 34) for k in range(1, 34)]
    outs, Cref, S, Ke, offC = _run_batched(dev, shapes, ta, tb, ls.FP32, 1000 + 2 * ta + tb)
    C = outs[0][:-1]
    assert np.isnan(outs[0][-1]), "wrote past the last C"
    bad, st = oracle.check_kv(C, Cref, S, Ke, U32)
    assert bad == 0, st


@pytest.mark.parametrize("ta,tb", TRANS)
def test_tiny_shapes_single_calls(dev, ta, tb):
    rng = np.random.default_rng(5 + 2 * ta + tb)
    r = oracle.Drand48(77 + ta + 2 * tb)
    for _ in range(60):
        M, N, K = (int(x) for x in rng.integers(1, 40, size=3))
        A = r.matrix(*oracle.a_shape(ta, M, K))
        B = r.matrix(*oracle.b_shape(tb, N, K))
        dC = nan_c(M, N, dev)
        ls.sgemm(ta, tb, dC, dev_tensor(A, dev), dev_tensor(B, dev), M, N, K)
        bad, st = check_f32(dC.cpu().numpy(), A, B, ta, tb, M, N, K)
        assert bad == 0, ((M, N, K), st)


def test_all_ones_known_answer(dev):
    for (M, N, K) in [(80, 80, 80), (33, 17, 129), (1, 300, 5)]:
        for prec in (ls.FP32, ls.BF16X3, ls.BF16, ls.BF16X6):
            A = torch.ones(M * K, device=dev)
            B = torch.ones(K * N, device=dev)
            dC = nan_c(M, N, dev)
            ls.sgemm_ex(0, 0, dC, A, B, M, N, K, precision=prec)
            torch.cuda.synchronize()
            assert (dC == K).all(), (M, N, K, prec)


def test_k_zero_and_empty(dev):
    A = torch.zeros(1, device=dev)
    dC = nan_c(5, 7, dev)
    ls.sgemm(0, 0, dC, A, A, 5, 7, 0)
    assert (dC == 0).all()
    ls.sgemm(0, 0, dC, A, A, 0, 7, 3)  # empty: no-op
    dD = torch.full((4, 3), float("nan"), dtype=torch.float64, device=dev)
    ls.dgemm(1, 1, dD, A, A, 4, 3, 0)
    assert (dD == 0).all()


def _sample_rows_check(dC, dA, dB, ta, tb, M, N, K, prec, rng, nrows=48):
    """Oracle check of a sample of C's rows (gathered on the device)."""
    rows = np.unique(np.concatenate([[0, M - 1], rng.integers(0, M, size=nrows)]))
    ridx = torch.from_numpy(rows).to(dC.device)
    opA = dA.view(*oracle.a_shape(ta, M, K))
    Asub = (opA[:, ridx].T if ta else opA[ridx]).contiguous().cpu().numpy()
    C = dC.view(M, N)[ridx].cpu().numpy()
    return check_f32(C, Asub, dB.cpu().numpy(), 0, tb, len(rows), N, K, prec)


def _full_check_f64(dC, dA, dB, ta, tb, M, N, K, prec):
    """Every element of C against a float64 product of the same operands
    computed on the device (its own error, ~K 2^-53 |A||B|, is negligible
    against the bound): |C - C64| <= (2 K u + extra) |A||B|."""
    A = dA.view(*oracle.a_shape(ta, M, K)).double()
    B = dB.view(*oracle.b_shape(tb, N, K)).double()
    A = A.T.contiguous() if ta else A
    B = B.T.contiguous() if tb else B
    ref = A @ B
    den = A.abs() @ B.abs()
    err = (dC.view(M, N).double() - ref).abs()
    bound = den * (2.0 * max(K, 1) * U32 + PREC_EXTRA[prec])
    return int((~(err <= bound)).sum()), float((err / den.clamp_min(1e-300)).max())


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
@pytest.mark.parametrize("tb", [0, 1])
def test_mp_tall_skinny(dev, prec, tb):
    """Irregular-shaped sgemm_mp: M in {4096, 16384, 65536} x N in {16, 32, 64}
    x K in {64, 256, 1024}; NN and NT; rows sampled for the oracle, every
    element of C against a float64 product on the device, and NaN prefill
    checks every element is written."""
    rng = np.random.default_rng(11 + tb)
    for M in (4096, 16384, 65536):
        for N in (16, 32, 64):
            for K in (64, 256, 1024):
                gen = torch.Generator(device=dev).manual_seed(M + N + K)
                dA = torch.rand(M * K, device=dev, generator=gen) * 2 - 1
                dB = torch.rand(K * N, device=dev, generator=gen) * 2 - 1
                dC = nan_c(M, N, dev)
                ls.sgemm_mp_ex(0, tb, dC, dA, dB, M, N, K, precision=prec)
                torch.cuda.synchronize()
                assert bool(torch.isfinite(dC).all()), (M, N, K)
                bad, st = _sample_rows_check(dC, dA, dB, 0, tb, M, N, K, prec, rng)
                assert bad == 0, ((M, N, K), st)
                bad, mx = _full_check_f64(dC, dA, dB, 0, tb, M, N, K, prec)
                assert bad == 0, ((M, N, K), bad, mx)


@pytest.mark.parametrize("ta,tb", TRANS)
def test_mp_short_wide_and_transposes(dev, ta, tb):
    rng = np.random.default_rng(3)
    for (M, N, K) in [(16, 4096, 256), (64, 65536, 64), (5, 10000, 33), (256, 2048, 500), (333, 77, 1500)]:
        gen = torch.Generator(device=dev).manual_seed(M * 7 + N + K)
        dA = torch.rand(M * K, device=dev, generator=gen) * 2 - 1
        dB = torch.rand(K * N, device=dev, generator=gen) * 2 - 1
        dC = nan_c(M, N, dev)
        ls.sgemm_mp(ta, tb, dC, dA, dB, M, N, K)
        assert bool(torch.isfinite(dC).all()), (M, N, K)
        bad, st = _sample_rows_check(dC, dA, dB, ta, tb, M, N, K, ls.FP32, rng)
        assert bad == 0, ((M, N, K), st)
        bad, mx = _full_check_f64(dC, dA, dB, ta, tb, M, N, K, ls.FP32)
        assert bad == 0, ((M, N, K), bad, mx)


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X6])
@pytest.mark.parametrize("K", [1, 2, 3])
def test_fp32_bound_at_tiny_k(dev, prec, K):
    """The FP32 bound is tightest at K = 1 (2 u |a||b| per element, where a
    single rounding may already take u): 640 x 640 x K products, every
    element held to 2*K*u.  (A three-piece split that dropped its two
    products of size ~2^-24 |a||b| reached 1.33 of the bound here.)"""
    r = oracle.Drand48(7 + K)
    M = N = 640
    A = r.matrix(M, K)
    B = r.matrix(K, N)
    dC = nan_c(M, N, dev)
    ls.sgemm_ex(0, 0, dC, dev_tensor(A, dev), dev_tensor(B, dev), M, N, K, prec)
    bad, st = check_f32(dC.cpu().numpy(), A, B, 0, 0, M, N, K, prec)
    assert bad == 0, st


def test_split_k_after_graph_launch(dev):
    """Split-K GEMMs keep a tile's K slices on one XCD; the dispatcher's
    round-robin start shifts after a graph with forked branches has run (a
    batched plan), so run one first, then split-K GEMMs in both float modes,
    C prefilled with NaN, every element against the oracle."""
    shapes = [(m, n, k) for m in (9, 70, 130) for n in (5, 40, 100) for k in (3, 50, 120)]
    _run_batched(dev, shapes, 0, 0, ls.FP32, seed=11)
    for prec in (ls.FP32, ls.BF16X3):
        # (GEMMs with a short side and a long one of 1024 or more run on the
        # short-side kernel, which does not split K)
        for (M, N, K, tb) in [(1000, 16, 1024, 0), (16, 1000, 1024, 1), (900, 48, 768, 0)]:
            assert ls.query_plan(prec, M, N, K)["ksplit"] > 1
            r = oracle.Drand48(M + K)
            A = r.matrix(M, K)
            B = r.matrix(*oracle.b_shape(tb, N, K))
            dC = nan_c(M, N, dev)
            ls.sgemm_mp_ex(0, tb, dC, dev_tensor(A, dev), dev_tensor(B, dev), M, N, K, prec)
            bad, st = check_f32(dC.cpu().numpy(), A, B, 0, tb, M, N, K, prec)
            assert bad == 0, (prec, M, N, K, st)


@pytest.mark.parametrize("ta,tb", TRANS)
def test_dgemm(dev, ta, tb):
    """CP2K block shapes (benchmark/CP2K_kernels:74-76) plus the M % 7 in
    {1, 3, 4} rows the reference's 7x6 kernel never writes (M = 4, 22, 32)."""
    r = oracle.Drand48(2024 + ta + 2 * tb)
    shapes = [(5, 5, 5), (13, 5, 13), (13, 13, 13), (23, 23, 23), (26, 26, 13), (4, 9, 16), (22, 4, 32),
              (32, 32, 32), (1, 1, 1), (100, 70, 300), (513, 129, 67)]
    for (M, N, K) in shapes:
        A = r.matrix_d(*oracle.a_shape(ta, M, K))
        B = r.matrix_d(*oracle.b_shape(tb, N, K))
        dC = nan_c(M, N, dev, torch.float64)
        ls.dgemm(ta, tb, dC, dev_tensor(A, dev), dev_tensor(B, dev), M, N, K)
        Cref, absAB = oracle.dgemm_ref(ta, tb, A, B, M, N, K)
        bad, st = oracle.check(dC.cpu().numpy(), Cref, absAB, K, U64)
        assert bad == 0, ((M, N, K), st)


@pytest.mark.parametrize("prec", [ls.BF16X3, ls.BF16, ls.BF16X6])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_precision_modes(dev, prec, ta, tb):
    r = oracle.Drand48(31 + prec)
    for (M, N, K) in [(80, 80, 80), (128, 128, 128), (37, 91, 200), (7, 5, 3), (200, 64, 1000)]:
        A = r.matrix(*oracle.a_shape(ta, M, K))
        B = r.matrix(*oracle.b_shape(tb, N, K))
        dC = nan_c(M, N, dev)
        ls.sgemm_ex(ta, tb, dC, dev_tensor(A, dev), dev_tensor(B, dev), M, N, K, precision=prec)
        torch.cuda.synchronize()
        bad, st = check_f32(dC.cpu().numpy(), A, B, ta, tb, M, N, K, prec)
        assert bad == 0, ((M, N, K), st)


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3, ls.BF16X6, ls.BF16])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_batched_uneven(dev, prec, ta, tb):
    """Uneven batch, M, N, K ~ U{1..128}; replayed twice: identical results.
    Every tile bucket of the persistent kernels (their split modes with the
    interleaved placement and, in BF16X6's 128 x 128 bucket, one LDS buffer)."""
    rng = np.random.default_rng(17 + prec)
    shapes = [tuple(int(x) for x in rng.integers(1, 129, size=3)) for _ in range(400)]
    outs, Cref, S, Ke, offC = _run_batched(dev, shapes, ta, tb, prec, 555, replays=2)
    assert np.array_equal(outs[0], outs[1], equal_nan=True)
    bad, st = oracle.check_kv(outs[0][:-1], Cref, S, Ke, U32, PREC_EXTRA[prec])
    assert bad == 0, st
    assert np.isnan(outs[0][-1])


def test_dgemm_batched_cp2k(dev):
    """FP64 batch with M, N, K drawn from {4,5,9,13,16,22,23,26,32}."""
    dims = np.array([4, 5, 9, 13, 16, 22, 23, 26, 32])
    rng = np.random.default_rng(4)
    shapes = [tuple(int(x) for x in rng.choice(dims, size=3)) for _ in range(2000)]
    outs, Cref, S, Ke, offC = _run_batched(dev, shapes, 0, 0, ls.FP64, 99, f64=True)
    bad, st = oracle.check_kv(outs[0][:-1], Cref, S, Ke, U64)
    assert bad == 0, st


def test_one_shot_batched_matches_plan(dev):
    rng = np.random.default_rng(8)
    shapes = [tuple(int(x) for x in rng.integers(1, 70, size=3)) for _ in range(50)]
    A, B, offA, offB, offC = _flat_batch(shapes, 0, 0, 8)
    dA, dB = dev_tensor(A, dev), dev_tensor(B, dev)
    dC = torch.full((int(offC[-1]),), float("nan"), device=dev)
    Cs = [dC.data_ptr() + 4 * int(o) for o in offC[:-1]]
    As = [dA.data_ptr() + 4 * int(o) for o in offA[:-1]]
    Bs = [dB.data_ptr() + 4 * int(o) for o in offB[:-1]]
    ls.sgemm_batched(0, 0, Cs, As, Bs, *zip(*shapes))
    Cref, S, Ke = oracle.gemm_ref_flat(0, 0, A, B, offA[:-1], offB[:-1], offC[:-1], *zip(*shapes), int(offC[-1]))
    assert oracle.check_kv(dC.cpu().numpy(), Cref, S, Ke, U32)[0] == 0


def test_determinism_repeated_calls(dev):
    r = oracle.Drand48(1)
    M, N, K = 4096, 32, 1024
    A = dev_tensor(r.matrix(M, K), dev)
    B = dev_tensor(r.matrix(K, N), dev)
    C1 = nan_c(M, N, dev)
    C2 = nan_c(M, N, dev)
    ls.sgemm_mp(0, 0, C1, A, B, M, N, K)
    ls.sgemm_mp(0, 0, C2, A, B, M, N, K)
    assert torch.equal(C1, C2)


@pytest.mark.parametrize("ta,tb", TRANS)
def test_backward_matches_matmul(dev, ta, tb):
    torch.manual_seed(3)
    for (M, N, K) in [(64, 300, 72), (17, 33, 5), (128, 49, 256)]:
        A = torch.randn(*oracle.a_shape(ta, M, K), dtype=torch.float64, device=dev).float()
        B = torch.randn(*oracle.b_shape(tb, N, K), dtype=torch.float64, device=dev).float()
        dC = torch.randn(M, N, device=dev)
        dA = torch.full_like(A, float("nan"))
        dB = torch.full_like(B, float("nan"))
        ls.sgemm_backward(ta, tb, dC, A, B, dA, dB, M, N, K)
        torch.cuda.synchronize()
        A64, B64, dC64 = A.double(), B.double(), dC.double()
        opA = A64.T if ta else A64
        opB = B64.T if tb else B64
        gA = dC64 @ opB.T
        gB = opA.T @ dC64
        refA = gA.T if ta else gA
        refB = gB.T if tb else gB
        bndA = (dC64.abs() @ opB.abs().T)
        bndB = (opA.abs().T @ dC64.abs())
        bndA = (bndA.T if ta else bndA) * 2 * N * U32
        bndB = (bndB.T if tb else bndB) * 2 * M * U32
        assert ((dA.double() - refA).abs() <= bndA).all()
        assert ((dB.double() - refB).abs() <= bndB).all()


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
def test_autotuned_plan(dev, prec):
    """Tune the tile per shape class on an uneven batch, then build the plan:
    every GEMM still meets the bound, and the tiles used are candidates."""
    from libshalom_amd.workloads import uniform_shapes
    shapes = uniform_shapes(3000, 1, 128, seed=5)
    Ms, Ns, Ks = (np.array(x, dtype=np.int64) for x in zip(*shapes))
    try:
        n = ls.autotune_batched(prec, 0, 0, Ms, Ns, Ks)
        assert n > 0
        tiles = {ls.query_tuned_tile(prec, 0, 0, *s) for s in shapes}
        assert tiles <= {(32, 32), (64, 64), (128, 128)}
        outs, Cref, S, Ke, offC = _run_batched(dev, shapes[:600], 0, 0, prec, seed=9)
        bad, st = oracle.check_kv(outs[0][:-1], Cref, S, Ke, U32, PREC_EXTRA[prec])
        assert bad == 0, st
    finally:
        ls.autotune_reset()


@pytest.mark.parametrize("prec", [ls.BF16X3, ls.BF16])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_batched_large_tiles_all_trans(dev, prec, ta, tb):
    """Batched GEMMs with max(M, N) > 64 and M, N, K <= 128 (the 128 x 128
    bucket) in the BF16 modes and all four trans combinations, incl. K = 1,
    K = 128 and N or M = 1, back to back in memory: every element within the
    bound, replays identical, nothing written past a block."""
    rng = np.random.default_rng(40 + 2 * ta + tb + 7 * prec)
    shapes = [tuple(int(x) for x in rng.integers(1, 129, size=3)) for _ in range(300)]
    shapes += [(128, 128, 128), (128, 1, 128), (1, 128, 65), (65, 70, 1), (100, 128, 31), (97, 3, 128)]
    shapes = [s for s in shapes if max(s[0], s[1]) > 64]
    outs, Cref, S, Ke, offC = _run_batched(dev, shapes, ta, tb, prec, seed=77, replays=2)
    assert np.array_equal(outs[0], outs[1], equal_nan=True)
    bad, st = oracle.check_kv(outs[0][:-1], Cref, S, Ke, U32, PREC_EXTRA[prec])
    assert bad == 0, st
    assert np.isnan(outs[0][-1])


@pytest.mark.parametrize("f64", [False, True])
def test_plan_replays_with_eager_work_between(dev, f64):
    """A plan's graph runs its buckets on parallel branches.  Replay it four
    times with unrelated eager work on the same stream in between (a single
    GEMM and a fill of C), and check every replay against the oracle and bit
    for bit against the first."""
    rng = np.random.default_rng(31 + int(f64))
    if f64:
        dims = [4, 5, 9, 13, 16, 22, 23, 26, 32]  # 16x16 and 32x32 buckets
        shapes = [tuple(int(rng.choice(dims)) for _ in range(3)) for _ in range(400)]
    else:
        shapes = [tuple(int(x) for x in rng.integers(1, 129, size=3)) for _ in range(400)]  # every bucket
    prec = ls.FP32
    A, B, offA, offB, offC = _flat_batch(shapes, 0, 0, 17, f64)
    dt = torch.float64 if f64 else torch.float32
    dA, dB = dev_tensor(A, dev), dev_tensor(B, dev)
    dC = torch.full((int(offC[-1]),), float("nan"), dtype=dt, device=dev)
    es = dA.element_size()
    Cs = [dC.data_ptr() + es * int(o) for o in offC[:-1]]
    As = [dA.data_ptr() + es * int(o) for o in offA[:-1]]
    Bs = [dB.data_ptr() + es * int(o) for o in offB[:-1]]
    Ms, Ns, Ks = zip(*shapes)
    Cref, S, Ke = oracle.gemm_ref_flat(0, 0, A, B, offA[:-1], offB[:-1], offC[:-1], Ms, Ns, Ks, int(offC[-1]))
    X = torch.randn(300, 70, device=dev)
    Y = torch.randn(70, 90, device=dev)
    Z = torch.empty(300, 90, device=dev)
    first = None
    with ls.BatchedPlan(0, 0, Cs, As, Bs, Ms, Ns, Ks, precision=prec, fp64=f64) as plan:
        for rep in range(4):
            ls.sgemm_ex(0, 0, Z, X, Y, 300, 90, 70, ls.FP32)  # eager work on the stream
            dC.fill_(float("nan"))
            plan.execute()
            if rep % 2:
                torch.cuda.synchronize()
            out = dC.cpu().numpy().copy()
            bad, st = oracle.check_kv(out, Cref, S, Ke, 2.0 ** -53 if f64 else U32)
            assert bad == 0, (rep, st)
            if first is None:
                first = out
            assert np.array_equal(out, first), rep
    ref = X.double() @ Y.double()
    assert ((Z.double() - ref).abs() <= 2 * 70 * U32 * (X.double().abs() @ Y.double().abs())).all()
