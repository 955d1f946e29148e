"""Host-side tests of the C-ABI library (no GPU): it loads, exports exactly
what include/LibShalom.h declares, rejects bad arguments before touching the
device, and its planner obeys its invariants."""
import ctypes as ct
import os
import re

import pytest

import libshalom_amd as ls


def _declared():
    with open(ls.HEADER_PATH) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(LibShalom_\w+)\s*\(", text)))


def test_header_symbols_exported():
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(ls.lib(), n), f"{n} declared but not exported"
        assert n in ls.SIGNATURES, f"{n} has no ctypes signature"
    assert sorted(ls.SIGNATURES) == names


def test_reference_api_signatures_match_reference_order():
    # C first, then A, B, then M, N, K (NN_LIB/LibShalom.h:22-39)
    for n in ("LibShalom_sgemm", "LibShalom_sgemm_mp", "LibShalom_dgemm"):
        res, args = ls.SIGNATURES[n]
        assert res is None and len(args) == 8


def test_libsmm_alias():
    import os
    alias = os.path.join(os.path.dirname(ls.LIB_PATH), "libsmm.so")
    assert os.path.exists(alias)
    ct.CDLL(alias).LibShalom_sgemm  # noqa: B018


@pytest.mark.parametrize("args,code", [
    ((0, 0, 0x1000, 0x1000, 0x1000, -1, 4, 4), 1),
    ((0, 0, 0x1000, 0x1000, 0x1000, 4, -4, 4), 1),
    ((2, 0, 0x1000, 0x1000, 0x1000, 4, 4, 4), 1),
    ((0, 7, 0x1000, 0x1000, 0x1000, 4, 4, 4), 1),
    ((0, 0, None, 0x1000, 0x1000, 4, 4, 4), 2),
    ((0, 0, 0x1000, None, 0x1000, 4, 4, 4), 2),
])
def test_argument_errors_before_device(args, code):
    L = ls.lib()
    L.LibShalom_sgemm(*args)
    assert L.LibShalom_get_last_error() == code
    L.LibShalom_dgemm(*args)
    assert L.LibShalom_get_last_error() == code
    with pytest.raises(ls.LibShalomError) as e:
        ls.sgemm_ex(*args, precision=ls.FP32, stream=0)
    assert e.value.status == code


@pytest.mark.parametrize("kw,code", [
    (dict(batch=-1), 1),
    (dict(strideA=-1), 1),
    (dict(strideC=-5), 1),
    (dict(M=-1), 1),
    (dict(transa=3), 1),
    (dict(precision=ls.FP64), 1),
    (dict(dC=None), 2),
    (dict(dA=None), 2),
    (dict(B=None), 2),
])
def test_backward_argument_errors_before_device(kw, code):
    args = dict(transa=0, transb=0, dC=0x1000, A=0x1000, B=0x1000, dA=0x1000, dB=0x1000, M=4, N=4, K=4,
                batch=2, strideA=0, strideB=16, strideC=16, precision=ls.FP32)
    args.update(kw)
    L = ls.lib()
    st = L.LibShalom_sgemm_backward_strided(args["transa"], args["transb"], args["dC"], args["A"], args["B"],
                                            args["dA"], args["dB"], args["M"], args["N"], args["K"],
                                            args["batch"], args["strideA"], args["strideB"], args["strideC"],
                                            args["precision"], None)
    assert st == code and L.LibShalom_get_last_error() == code
    if args["batch"] == 2 and args["strideA"] == 0 and args["strideC"] == 16:
        st = L.LibShalom_sgemm_backward(args["transa"], args["transb"], args["dC"], args["A"], args["B"],
                                        args["dA"], args["dB"], args["M"], args["N"], args["K"],
                                        args["precision"], None)
        assert st == code


@pytest.mark.parametrize("ta,tb,M,N,K,lda,ldb,ldc,code", [
    (0, 0, 8, 6, 5, 4, 6, 6, 1),      # lda < K
    (1, 0, 8, 6, 5, 7, 6, 6, 1),      # transposed A: lda < M
    (0, 0, 8, 6, 5, 5, 5, 6, 1),      # ldb < N
    (0, 1, 8, 6, 5, 5, 4, 6, 1),      # transposed B: ldb < K
    (0, 0, 8, 6, 5, 5, 6, 5, 1),      # ldc < N
    (0, 0, 8, 6, 0, 0, 6, 6, 1),      # K = 0 still needs lda >= 1
    (0, 0, 8, 6, 5, 1 << 21, 6, 6, 5),  # offsets would overflow: not supported
    (2, 0, 8, 6, 5, 5, 6, 6, 1),
])
def test_general_argument_errors_before_device(ta, tb, M, N, K, lda, ldb, ldc, code):
    L = ls.lib()
    st = L.LibShalom_sgemm_general(ta, tb, M, N, K, 1.0, 0x1000, lda, 0x1000, ldb, 0.0, 0x1000, ldc,
                                   ls.FP32, None)
    assert st == code and L.LibShalom_get_last_error() == code
    st = L.LibShalom_dgemm_general(ta, tb, M, N, K, 1.0, 0x1000, lda, 0x1000, ldb, 0.0, 0x1000, ldc, None)
    assert st == code
    with pytest.raises(ls.LibShalomError) as e:
        ls.sgemm_general(ta, tb, M, N, K, 2.0, 0x1000, lda, 0x1000, ldb, 1.0, 0x1000, ldc, ls.BF16X3, 0)
    assert e.value.status == code


def test_general_null_pointers_rejected():
    L = ls.lib()
    assert L.LibShalom_sgemm_general(0, 0, 4, 4, 4, 1.0, None, 4, 0x1000, 4, 0.0, 0x1000, 4, ls.FP32, None) == 2
    assert L.LibShalom_dgemm_general(0, 0, 4, 4, 4, 1.0, 0x1000, 4, 0x1000, 4, 0.0, None, 4, None) == 2
    assert L.LibShalom_sgemm_general(0, 0, 4, 4, 4, 1.0, 0x1000, 4, 0x1000, 4, 0.0, 0x1000, 4, 9, None) == 1


def test_general_alpha_zero_needs_no_operands():
    """BLAS leaves A and B unreferenced when alpha is 0: null A / B pass the
    argument checks (and the call then stops at the missing device)."""
    L = ls.lib()
    assert L.LibShalom_sgemm_general(0, 0, 4, 4, 4, 0.0, None, 4, None, 4, 1.0, 0x1000, 4, ls.FP32, None) == 6
    assert L.LibShalom_dgemm_general(0, 0, 4, 4, 4, 0.0, None, 4, None, 4, 0.0, 0x1000, 4, None) == 6
    assert L.LibShalom_sgemm_general(0, 0, 4, 4, 4, 0.0, None, 4, None, 4, 1.0, None, 4, ls.FP32, None) == 2


def test_backward_empty_batch_is_a_no_op():
    L = ls.lib()
    assert L.LibShalom_sgemm_backward_strided(0, 0, None, None, None, None, None, 4, 4, 4, 0, 0, 0, 0,
                                              ls.FP32, None) == 0


def test_backward_transpose_tables_match_derivation():
    """The (TA, TB) that the dA and dB problems use, restated from the chain
    rule: dA = dC op(B)^T (stored transposed when transa), dB = op(A)^T dC
    (stored transposed when transb); checked with explicit matrices."""
    import numpy as np
    rng = np.random.default_rng(0)
    M, N, K = 3, 4, 5
    for ta in (0, 1):
        for tb in (0, 1):
            A = rng.standard_normal((K, M) if ta else (M, K))
            B = rng.standard_normal((N, K) if tb else (K, N))
            dC = rng.standard_normal((M, N))
            opA = A.T if ta else A
            opB = B.T if tb else B
            # finite-difference-free check of the closed forms on a bilinear loss
            dA = dC @ opB.T
            dB = opA.T @ dC
            dA = dA.T if ta else dA
            dB = dB.T if tb else dB
            eps = 1e-6
            E = np.zeros_like(A)
            E[0, 1] = 1.0
            Ap = A + eps * E
            opAp = Ap.T if ta else Ap
            fd = (np.sum(dC * (opAp @ opB)) - np.sum(dC * (opA @ opB))) / eps
            assert abs(fd - dA[0, 1]) < 1e-5
            E = np.zeros_like(B)
            E[1, 0] = 1.0
            Bp = B + eps * E
            opBp = Bp.T if tb else Bp
            fd = (np.sum(dC * (opA @ opBp)) - np.sum(dC * (opA @ opB))) / eps
            assert abs(fd - dB[1, 0]) < 1e-5


def test_bad_precision_rejected():
    with pytest.raises(ls.LibShalomError):
        ls.set_precision(9)
    with pytest.raises(ls.LibShalomError):
        ls.sgemm_ex(0, 0, 0x1000, 0x1000, 0x1000, 4, 4, 4, precision=4, stream=0)
    ls.set_precision(ls.BF16X3)
    assert ls.get_precision() == ls.BF16X3
    ls.set_precision(ls.FP32)
    assert ls.get_precision() == ls.FP32


def test_thread_config_is_recorded_noop():
    ls.set_thread_nums(64)
    assert ls.get_thread_config()[0] == 64
    ls.set_thread_ways(8, 4)
    assert ls.get_thread_config() == (32, 8, 4)


def test_status_strings():
    for s in range(7):
        assert ls.lib().LibShalom_status_string(s)


def test_batched_plan_rejects_bad_input():
    h = ct.c_void_p()
    L = ls.lib()
    # negative count
    assert L.LibShalom_sgemm_batched_plan(ct.byref(h), 0, 0, None, None, None, None, None, None, -1, 0) == 1
    # null arrays with count > 0
    assert L.LibShalom_sgemm_batched_plan(ct.byref(h), 0, 0, None, None, None, None, None, None, 3, 0) == 2
    # bad precision
    assert L.LibShalom_sgemm_batched_plan(ct.byref(h), 0, 0, None, None, None, None, None, None, 0, 11) == 1
    assert L.LibShalom_plan_execute(None, None) == 2
    assert L.LibShalom_plan_destroy(None) == 0


SHAPES = [(80, 80, 80), (8, 8, 8), (128, 128, 128), (4096, 16, 1024), (65536, 64, 1024), (16, 65536, 256),
          (64, 3136, 576), (512, 49, 4608), (1, 1, 1), (4096, 4096, 4096), (3, 100000, 7), (100000, 3, 7),
          (256, 2048, 5000), (31999, 1024, 88)]


TILES = {(32, 32), (64, 64), (128, 128), (128, 32), (128, 64), (32, 128)}


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3, ls.BF16, ls.BF16X6, ls.FP64])
@pytest.mark.parametrize("shape", SHAPES)
def test_planner_invariants(prec, shape):
    M, N, K = shape
    for mp in (False, True):
        p = ls.query_plan(prec, M, N, K, num_cu=256, mp=mp)
        tm, tn = p["tile"]
        if (tm, tn) not in TILES:  # the short-side kernel: 16 long-side indices, K whole
            assert prec != ls.FP64 and ((N <= 64 and M >= 1024) or (M <= 64 and N >= 1024)), (M, N, K, p)
            assert (tm if N <= 64 else tn) == 16 and p["ksplit"] == 1 and p["kper"] == max(K, 1), (M, N, K, p)
            continue
        assert (tm, tn) in TILES
        if prec == ls.FP64:
            assert (tm, tn) != (128, 128)
        ks, kp = p["ksplit"], p["kper"]
        assert ks >= 1 and kp >= 1
        assert ks * kp >= K and (ks - 1) * kp < max(K, 1)
        if ks > 1:
            assert kp % (16 if prec == ls.FP64 else 32) == 0
            assert kp >= 4 * (16 if prec == ls.FP64 else 32)


def test_planner_fills_the_chip():
    # tall-skinny: few tiles, so the planner splits K
    p = ls.query_plan(ls.BF16X3, 4096, 16, 1024, num_cu=256, mp=True)
    tiles = -(-4096 // p["tile"][0]) * -(-16 // p["tile"][1])
    assert tiles * p["ksplit"] >= 64
    # big square: plenty of tiles, no split
    p = ls.query_plan(ls.FP32, 4096, 4096, 4096, num_cu=256)
    assert p["ksplit"] == 1


# The wave grid of each tile (TileCfg in gemm_core.hpp): 128 x 128 tiles run
# eight waves in a 4 x 2 grid; the other square tiles four waves in a 2 x 2
# grid; the tall tiles (128 x 32, 128 x 64) split their rows over four waves,
# the wide one (32 x 128) its columns.  In the split modes' tile kernels a
# wave owns a contiguous block of 16 x 16 sub-tiles; in FP32 / FP64 its
# sub-tile rows and columns are interleaved with the other waves' (row blocks
# g, g + GM, ...).
WAVE_GRID = {(32, 32): (2, 2), (64, 64): (2, 2), (128, 128): (4, 2), (128, 32): (4, 1),
             (128, 64): (4, 1), (32, 128): (1, 4)}


def idle_fraction(tile, M, N, prec):
    """Fraction of a tile's waves holding no part of an M x N GEMM (its first tile)."""
    bm, bn = tile
    gm, gn = WAVE_GRID[tile]
    mt, nt = -(-min(bm, M) // 16), -(-min(bn, N) // 16)
    if prec in (ls.FP32, ls.FP64):
        live_r, live_c = min(gm, mt), min(gn, nt)
    else:
        wm, wn = bm // 16 // gm, bn // 16 // gn
        live_r, live_c = min(gm, -(-mt // wm)), min(gn, -(-nt // wn))
    return 1.0 - live_r * live_c / (gm * gn)


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3, ls.BF16X6, ls.FP64])
def test_planner_keeps_waves_busy_on_skinny_grid(prec):
    """On the tall-skinny / short-wide sgemm_mp grid, the chosen tile never
    leaves more than half its waves idle when another tile would not."""
    from libshalom_amd.workloads import TALL_SKINNY
    shapes = list(TALL_SKINNY) + [(N, M, K) for (M, N, K) in TALL_SKINNY]
    for (M, N, K) in shapes:
        tile = ls.query_plan(prec, M, N, K, num_cu=256, mp=True)["tile"]
        if tile not in WAVE_GRID:
            continue  # the FP32 short-side kernel: its four waves split K (test below)
        if idle_fraction(tile, M, N, prec) > 0.5:
            others = [t for t in WAVE_GRID if not (prec == ls.FP64 and t == (128, 128))]
            assert all(idle_fraction(t, M, N, prec) > 0.5 for t in others), (M, N, K, tile)


def test_short_side_plans():
    """FP32 and split-mode GEMMs with a short side (N <= 64 and M >= 1024, or
    M <= 64 and N >= 1024) run on the short-side kernel: 16 long-side indices
    per workgroup, the short side in 16/32/64-wide blocks, K whole; FP64 and
    squarer shapes keep the tile planner."""
    for prec in (ls.FP32, ls.BF16X3, ls.BF16, ls.BF16X6):
        for (M, N, K, want) in [(4096, 16, 1024, (16, 16)), (4096, 17, 1024, (16, 32)), (8192, 16, 1024, (16, 16)),
                                (16, 4096, 1024, (16, 16)), (48, 4096, 1024, (64, 16)), (1024, 32, 1024, (16, 32)),
                                (4096, 16, 4096, (16, 16)), (4096, 64, 512, (16, 64))]:
            if prec != ls.FP32 and 64 in want:
                continue  # the split modes keep their tile plans on short sides above 32
            p = ls.query_plan(prec, M, N, K, num_cu=256, mp=True)
            assert p["tile"] == want and p["ksplit"] == 1 and p["kper"] == K, (prec, M, N, K, p)
    # FP64, squarer shapes, small K, long chains of K x short-side work per
    # CU, and split-mode short sides above 32 keep the tile planner
    for (prec, M, N, K) in [(ls.FP64, 4096, 16, 1024), (ls.FP32, 1000, 16, 1024), (ls.BF16X3, 4096, 64, 1024),
                            (ls.BF16X6, 48, 4096, 1024),
                            (ls.FP32, 4096, 65, 1024), (ls.FP32, 65, 4096, 1024), (ls.FP32, 4096, 16, 256),
                            (ls.FP32, 1024, 64, 4096), (ls.FP32, 2048, 32, 2048), (ls.FP32, 65536, 64, 1024),
                            (ls.FP32, 16, 16384, 1024)]:
        p = ls.query_plan(prec, M, N, K, num_cu=256, mp=True)
        assert p["tile"] in TILES or (prec == ls.FP64 and p["tile"] == (16, 16)), (prec, M, N, K, p)


def test_planner_errors():
    with pytest.raises(ls.LibShalomError):
        ls.query_plan(7, 4, 4, 4)
    with pytest.raises(ls.LibShalomError):
        ls.query_plan(ls.FP32, -1, 4, 4)
    with pytest.raises(ls.LibShalomError):
        ls.query_plan(ls.FP32, 4, 4, 4, num_cu=0)


@pytest.mark.parametrize("prec", [ls.FP32, ls.BF16X3, ls.FP64])
def test_batched_tile_rule(prec):
    assert ls.query_batched_tile(prec, 5, 32, 100) == (32, 32)
    # FP64: one-wave 16 x 16 tiles for GEMMs with M, N <= 16
    assert ls.query_batched_tile(prec, 13, 16, 100) == ((16, 16) if prec == ls.FP64 else (32, 32))
    assert ls.query_batched_tile(prec, 33, 10, 1) == (64, 64)
    big = ls.query_batched_tile(prec, 100, 128, 7)
    assert big == ((64, 64) if prec == ls.FP64 else (128, 128))


def test_tuned_tile_defaults_to_fixed_rule():
    ls.autotune_reset()
    for prec in (ls.FP32, ls.BF16X3, ls.FP64):
        for (M, N, K) in [(1, 1, 1), (17, 40, 9), (64, 64, 64), (65, 3, 128), (128, 128, 1), (300, 20, 77)]:
            assert ls.query_tuned_tile(prec, 0, 0, M, N, K) == ls.query_batched_tile(prec, M, N, K)


def test_autotune_argument_errors_before_device():
    L = ls.lib()
    assert L.LibShalom_autotune_batched(9, 0, 0, None, None, None, 0, None, None) == 1
    assert L.LibShalom_autotune_batched(0, 2, 0, None, None, None, 0, None, None) == 1
    assert L.LibShalom_autotune_batched(0, 0, 0, None, None, None, 5, None, None) == 2
    with pytest.raises(ls.LibShalomError):
        ls.query_tuned_tile(ls.FP32, 0, 0, -1, 4, 4)


def test_tune_file_is_loaded(tmp_path):
    """A table entry in LIBSHALOM_TUNE_FILE (precision transa transb class
    tile) overrides the fixed rule for that class, in a fresh process."""
    import subprocess
    import sys
    # class of (M, N, K) = (20, 20, 20): M, N in (16, 32], K in (16, 32]
    cls = (1 * 7 + 1) * 5 + 1
    path = tmp_path / "tune.txt"
    path.write_text(f"{ls.BF16X3} 0 1 {cls} 2\n")  # TileId 2 = 128 x 128
    code = ("import libshalom_amd as ls; print(ls.query_tuned_tile(ls.BF16X3, 0, 1, 20, 20, 20), "
            "ls.query_tuned_tile(ls.BF16X3, 0, 0, 20, 20, 20))")
    env = dict(os.environ, LIBSHALOM_TUNE_FILE=str(path))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, check=True,
                         cwd=os.path.dirname(os.path.dirname(ls.LIB_PATH)))
    assert out.stdout.split() == ["(128,", "128)", "(32,", "32)"]


def test_tune_file_ignores_tiles_without_a_persistent_kernel(tmp_path):
    """Lines naming a tile that has no persistent kernel for the mode (the
    rectangular tiles), an out-of-range class or mode are ignored; the valid
    line still applies."""
    import subprocess
    import sys
    cls = (1 * 7 + 1) * 5 + 1  # (20, 20, 20)
    path = tmp_path / "tune.txt"
    path.write_text(f"{ls.BF16X3} 0 0 {cls} 3\n"      # 128 x 32: no persistent kernel
                    f"{ls.BF16X3} 0 1 {cls} 5\n"      # 32 x 128: none either
                    f"{ls.FP32} 0 0 {cls} 4\n"        # 128 x 64: none
                    f"{ls.FP32} 1 1 -3 2\n"           # bad class
                    f"9 0 0 {cls} 2\n"                # bad mode
                    f"{ls.FP32} 0 1 {cls} 1\n")       # valid: 64 x 64
    code = ("import libshalom_amd as ls; print(ls.query_tuned_tile(ls.BF16X3, 0, 0, 20, 20, 20), "
            "ls.query_tuned_tile(ls.BF16X3, 0, 1, 20, 20, 20), ls.query_tuned_tile(ls.FP32, 0, 0, 20, 20, 20), "
            "ls.query_tuned_tile(ls.FP32, 0, 1, 20, 20, 20))")
    env = dict(os.environ, LIBSHALOM_TUNE_FILE=str(path))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, check=True,
                         cwd=os.path.dirname(os.path.dirname(ls.LIB_PATH)))
    assert out.stdout.split() == ["(32,", "32)", "(32,", "32)", "(32,", "32)", "(64,", "64)"]


def test_c_dropin_program_builds_and_links():
    """The C drop-in caller compiles as C99 against the header (-Wall -Wextra
    -Werror) and links with -lsmm; without a GPU its call reports
    LIBSHALOM_NO_DEVICE as its exit status."""
    import subprocess
    root = os.path.dirname(os.path.dirname(ls.LIB_PATH))
    cdir = os.path.join(root, "tests", "c")
    subprocess.run(["make", "-B", "-C", cdir], check=True, capture_output=True)
    r = subprocess.run([os.path.join(cdir, "dropin_hello"), os.devnull], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode in (0, 6), (r.returncode, r.stdout, r.stderr)
    assert "LibShalom_sgemm 80x80x80" in r.stdout


def test_host_layer_under_address_and_ub_sanitizers():
    """The host layer (api.cpp) rebuilt with -fsanitize=address,undefined
    (host code only) and driven from four threads through its argument
    checks, planner, tune-table lookups, plan builder and status bookkeeping
    (tests/c/host_asan.cpp): no sanitizer report, every status as expected."""
    import subprocess
    root = os.path.dirname(os.path.dirname(ls.LIB_PATH))
    cdir = os.path.join(root, "tests", "c")
    subprocess.run(["make", "-C", cdir, "host_asan"], check=True, capture_output=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(cdir, "host_asan")], capture_output=True, text=True, timeout=300, env=env,
                       cwd=cdir)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-4000:])
    assert "host_asan: 0 failures" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


# (batch, C, H, W, Cout, R, S, sh, sw, ph, pw, dh, dw, precision) -> status
CONV_ARG_CASES = [
    ((-1, 3, 8, 8, 4, 3, 3, 1, 1, 1, 1, 1, 1, 0), 1),   # negative batch
    ((1, 0, 8, 8, 4, 3, 3, 1, 1, 1, 1, 1, 1, 0), 1),    # no channels
    ((1, 3, 8, 8, 4, 3, 3, 0, 1, 1, 1, 1, 1, 0), 1),    # stride 0
    ((1, 3, 8, 8, 4, 3, 3, 1, 1, -1, 1, 1, 1, 0), 1),   # negative padding
    ((1, 3, 8, 8, 4, 3, 3, 1, 1, 1, 1, 0, 1, 0), 1),    # dilation 0
    ((1, 3, 2, 8, 4, 5, 3, 1, 1, 0, 1, 1, 1, 0), 1),    # filter taller than the input: P = 0
    ((1, 3, 8, 8, 4, 3, 3, 1, 1, 1, 1, 1, 1, 7), 1),    # bad precision
    ((1, 3, 8, 8, 4, 3, 3, 1, 1, 1, 1, 1, 1, 4), 1),    # FP64 has no convolution
    ((1, 1 << 20, 32, 32, 4, 3, 3, 1, 1, 1, 1, 1, 1, 0), 5),  # one image reaches 2 GiB
    ((1, 1, 1 << 28, 1, 1, 1, 1, 1, 1, 1 << 27, 0, 1, 1, 0), 5),  # H + 2 pad reaches 2^29
    ((1, 1, 1449, 1449, 1, 1, 1, 1, 1, 0, 0, 1, 1, 0), 5),      # P*Q reaches 2^21 (output ld)
    ((1, 233017, 1, 1, 1, 3, 3, 1, 1, 1, 1, 1, 1, 0), 5),       # C*R*S reaches 2^21 (filter ld)
    ((0, 3, 8, 8, 4, 3, 3, 1, 1, 1, 1, 1, 1, 0), 0),    # empty batch: nothing to do
]


@pytest.mark.parametrize("args,code", CONV_ARG_CASES)
def test_conv_argument_errors_before_device(args, code):
    L = ls.lib()
    st = L.LibShalom_sconv2d_forward(0x1000, 0x1000, 0x1000, *args, None)
    assert st == code and L.LibShalom_get_last_error() == code


def test_conv_null_pointers_rejected():
    L = ls.lib()
    g = (1, 3, 8, 8, 4, 3, 3, 1, 1, 1, 1, 1, 1, ls.FP32, None)
    assert L.LibShalom_sconv2d_forward(None, 0x1000, 0x1000, *g) == 2
    assert L.LibShalom_sconv2d_forward(0x1000, None, 0x1000, *g) == 2
    assert L.LibShalom_sconv2d_forward(0x1000, 0x1000, None, *g) == 2


@pytest.mark.parametrize("H,R,st,pad,dil", [(56, 3, 1, 1, 1), (7, 3, 1, 1, 1), (15, 3, 2, 1, 1), (15, 3, 1, 2, 2),
                                            (5, 5, 1, 3, 1), (9, 4, 2, 1, 2), (3, 5, 1, 0, 1), (10, 1, 3, 0, 1)])
def test_conv_output_size_matches_torch(H, R, st, pad, dil):
    import torch
    P, Q = ls.conv2d_output_size(H, H + 1, R, R, (st, st), (pad, pad), (dil, dil))
    x = torch.zeros(1, 1, H, H + 1)
    w = torch.zeros(1, 1, R, R)
    if H + 2 * pad < dil * (R - 1) + 1:
        assert P <= 0
        return
    y = torch.nn.functional.conv2d(x, w, stride=st, padding=pad, dilation=dil)
    assert (P, Q) == tuple(y.shape[2:])


def test_conv_backward_argument_errors_before_device():
    L = ls.lib()
    g = (1, 3, 8, 8, 4, 3, 3)
    # input gradient of a strided convolution: strides up to 8
    assert L.LibShalom_sconv2d_backward(0x1000, 0x1000, 0x1000, 0x1000, 0x1000, *g, 9, 1, 1, 1, 1, 1, 0, None) == 5
    assert L.LibShalom_sconv2d_backward(0x1000, 0x1000, 0x1000, 0x1000, 0x1000, *g, 1, 9, 1, 1, 1, 1, 0, None) == 5
    assert L.LibShalom_sconv2d_backward(0x1000, 0x1000, None, 0x1000, 0x1000, *g, 1, 1, 1, 1, 1, 1, 0, None) == 2
    assert L.LibShalom_sconv2d_backward(0x1000, None, 0x1000, None, 0x1000, *g, 1, 1, 1, 1, 1, 1, 0, None) == 2
    assert L.LibShalom_sconv2d_backward(None, 0x1000, 0x1000, 0x1000, None, *g, 1, 1, 1, 1, 1, 1, 0, None) == 2
    assert L.LibShalom_sconv2d_backward(None, None, 0x1000, 0x1000, 0x1000, *g, 1, 1, 1, 1, 1, 1, 0, None) == 0
    assert L.LibShalom_sconv2d_backward(0x1000, 0x1000, 0x1000, 0x1000, 0x1000, *g, 1, 1, 1, 1, 1, 1, 9, None) == 1
    assert L.LibShalom_sconv2d_backward(0x1000, 0x1000, 0x1000, 0x1000, 0x1000, 1, 3, 2, 8, 4, 5, 3,
                                        1, 1, 0, 1, 1, 1, 0, None) == 1  # P = 0
    assert L.LibShalom_sconv2d_backward(0x1000, 0x1000, 0x1000, 0x1000, 0x1000, 0, 3, 8, 8, 4, 3, 3,
                                        1, 1, 1, 1, 1, 1, 0, None) == 0  # empty batch


def test_conv_backward_leading_dimension_limits():
    L = ls.lib()
    P = 0x1000
    # Cout*R*S reaches 2^21 (the flipped filter's leading dimension)
    assert L.LibShalom_sconv2d_backward(P, P, P, P, P, 1, 1, 4, 4, 233017, 3, 3, 1, 1, 1, 1, 1, 1, 0, None) == 5
    # H*W reaches 2^21 (the input gradient's leading dimension)
    assert L.LibShalom_sconv2d_backward(P, P, P, P, P, 1, 1, 1449, 1449, 1, 1, 1, 1, 1, 0, 0, 1, 1, 0, None) == 5
