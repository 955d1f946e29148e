"""The host layer on the GPU: the C drop-in caller, host-operand staging,
workspaces under threads and graph capture, and BLAS's alpha = 0 rule.

* tests/c/dropin_hello is a C program compiled against include/LibShalom.h
  and linked with -lsmm (the reference's library name, NN_LIB/Makefile:7-8),
  passing malloc'd host buffers as the reference's harnesses do (Hello.c);
  its result is checked against the oracle's bound 2*K*u*(|A||B|).
* Host operands larger than the 4 MiB bounce buffers travel in chunks; the
  result must be bit-identical to the same call on device operands.
* Four host threads on the null stream, each with a split-K size of its own
  (so the per-stream fixup space grows while other threads hold it), every
  result checked against the oracle.
* A stream in graph capture cannot grow its buffers: the call is refused;
  after an eager call the capture succeeds, and a replay still computes the
  right product after a later, larger call has replaced the buffer.
* alpha = 0 reads neither A nor B (NaN-filled, or null pointers).
"""
import os
import subprocess
import threading

import numpy as np
import pytest
import torch

import libshalom_amd as ls
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
U32 = 2.0 ** -24
NOT_SUPPORTED = 5


def within(C, A, B, ta, tb, M, N, K):
    Cref, absAB = oracle.sgemm_ref(ta, tb, A, B, M, N, K)
    bad, st = oracle.check(C, Cref, absAB, K, oracle.U_F32)
    return bad, st


@pytest.mark.parametrize("mode,shape", [("hello", (80, 80, 80, 0, 1)), ("mp", (4096, 16, 1024, 0, 0))])
def test_c_dropin_program(dev, tmp_path, mode, shape):
    exe = os.path.join(ROOT, "tests", "c", "dropin_hello")
    if not os.path.exists(exe):  # normally built by __graft_entry__.build()
        subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "c")], check=True)
    out = tmp_path / "out.bin"
    r = subprocess.run([exe, str(out), mode], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    M, N, K, ta, tb = shape
    data = np.fromfile(out, dtype=np.float32)
    assert data.size == M * K + K * N + M * N
    A = data[:M * K]
    B = data[M * K:M * K + K * N]
    C = data[M * K + K * N:].reshape(M, N)
    bad, st = within(C, A, B, ta, tb, M, N, K)
    assert bad == 0, st
    assert np.isfinite(C).all()


@pytest.mark.parametrize("which", ["all", "A", "B", "C"])
def test_host_staging_in_chunks(dev, which):
    """A 12 MiB, B 6 MiB and C 18 MiB cross the 4 MiB bounce buffers (and A
    and B share a chunk); host and device operands mixed as `which` says."""
    M, N, K = 3000, 1500, 1024
    rng = np.random.default_rng(7)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    dA, dB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    dC = torch.full((M, N), float("nan"), device=dev)
    ls.sgemm_mp(0, 0, dC, dA, dB, M, N, K)
    want = dC.cpu().numpy()
    a = A if which in ("all", "A") else dA
    b = B if which in ("all", "B") else dB
    if which in ("all", "C"):
        C = np.full((M, N), np.nan, dtype=np.float32)
        ls.sgemm_mp(0, 0, C, a, b, M, N, K)
    else:
        dC2 = torch.full((M, N), float("nan"), device=dev)
        ls.sgemm_mp(0, 0, dC2, a, b, M, N, K)
        C = dC2.cpu().numpy()
    assert np.array_equal(C, want)
    rows = np.arange(0, M, 97)
    bad, st = within(C[rows], A[rows], B, 0, 0, len(rows), N, K)
    assert bad == 0, st


def test_reference_api_four_threads(dev):
    """Four threads on the null stream with different split-K sizes."""
    # split-K shapes of different partial-tile sizes (M < 1024 keeps them off
    # the FP32 short-side kernel, which needs no fixup space)
    shapes = [(1000, 16, 2048), (900, 32, 1024), (512, 16, 4096), (768, 40, 3000)]
    assert len({ls.query_plan(ls.FP32, *s, mp=True)["ksplit"] * s[0] * s[1] for s in shapes}) > 1
    inputs, refs = [], []
    for t, (M, N, K) in enumerate(shapes):
        rng = np.random.default_rng(40 + t)
        A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
        B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        inputs.append((A, B))
        refs.append(oracle.sgemm_ref(0, 0, A, B, M, N, K))
    dev_in = [(torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)) for A, B in inputs]
    torch.cuda.synchronize()
    errors = []

    def worker(t):
        M, N, K = shapes[t]
        dA, dB = dev_in[t]
        Cref, absAB = refs[t]
        dC = torch.empty((M, N), device=dev)
        for it in range(8):
            ls.sgemm_mp(0, 0, dC, dA, dB, M, N, K)  # synchronous, null stream
            C = dC.cpu().numpy()
            bad, st = oracle.check(C, Cref, absAB, K, oracle.U_F32)
            if bad:
                errors.append((t, it, bad, st))
                return
            dC.fill_(float("nan"))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors


def test_capture_refuses_growth_then_replays_after_growth(dev):
    # shapes on the tile kernels' split-K path (not short-sided: FP32 GEMMs
    # with N or M <= 64 take the short-side kernel, which needs no space)
    M, N, K = 128, 128, 8192
    assert ls.query_plan(ls.FP32, M, N, K, mp=True)["ksplit"] > 1
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    dA, dB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    dC = torch.full((M, N), float("nan"), device=dev)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    # a fresh stream has no fixup space: the captured call would allocate it
    g = torch.cuda.CUDAGraph()
    with pytest.raises(ls.LibShalomError) as ei:
        with torch.cuda.graph(g, stream=s):
            ls.sgemm_mp_ex(0, 0, dC, dA, dB, M, N, K, ls.FP32, s)
    assert ei.value.status == NOT_SUPPORTED
    torch.cuda.synchronize()
    # an eager call sizes the stream's space; the capture then succeeds
    ls.sgemm_mp_ex(0, 0, dC, dA, dB, M, N, K, ls.FP32, s)
    s.synchronize()
    want = dC.cpu().numpy()
    bad, st = within(want, A, B, 0, 0, M, N, K)
    assert bad == 0, st
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        ls.sgemm_mp_ex(0, 0, dC, dA, dB, M, N, K, ls.FP32, s)
    # a larger eager call on the same stream replaces (and keeps) the buffer
    M2, N2, K2 = 512, 512, 16384
    assert ls.query_plan(ls.FP32, M2, N2, K2, mp=True)["ksplit"] > 1
    dA2 = torch.rand(M2, K2, device=dev)
    dB2 = torch.rand(K2, N2, device=dev)
    dC2 = torch.empty(M2, N2, device=dev)
    for rep in range(3):
        dC.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(dC.cpu().numpy(), want), rep
        with torch.cuda.stream(s):
            ls.sgemm_mp_ex(0, 0, dC2, dA2, dB2, M2, N2, K2, ls.FP32, s)
        s.synchronize()


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3])
def test_alpha_zero_reads_neither_operand(dev, prec):
    M, N, K = 96, 80, 1024
    A = torch.full((M, K), float("nan"), device=dev)
    B = torch.full((K, N), float("inf"), device=dev)
    C0 = torch.rand(M, N, device=dev)
    C = C0.clone()
    ls.sgemm_general(0, 0, M, N, K, 0.0, A, K, B, N, 0.5, C, N, prec)
    torch.cuda.synchronize()
    assert torch.equal(C, C0 * 0.5)
    C.fill_(float("nan"))
    ls.sgemm_general(0, 0, M, N, K, 0.0, A, K, B, N, 0.0, C, N, prec)
    torch.cuda.synchronize()
    assert torch.equal(C, torch.zeros_like(C))
    # null operands are accepted when alpha is 0
    C = C0.clone()
    ls.sgemm_general(0, 0, M, N, K, 0.0, None, K, None, N, 2.0, C, N, prec)
    torch.cuda.synchronize()
    assert torch.equal(C, C0 * 2.0)
    Cd = C0.double()
    ls.dgemm_general(0, 0, M, N, K, 0.0, None, K, None, N, -1.0, Cd, N)
    torch.cuda.synchronize()
    assert torch.equal(Cd, -C0.double())
