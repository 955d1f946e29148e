// gemm_core.hpp -- the tile GEMM at the heart of every LibShalom entry point,
// written for gfx950 (CDNA4): 64-wide wavefronts, MFMA matrix cores, LDS.
//
// One workgroup (8 waves for 128 x 128 tiles, 4 otherwise) computes one
// BM x BN tile of C = op(A) * op(B) (row-major, beta = 0 overwrite) over a K
// range.
//
// * Fused pack.  LibShalom packs operand panels inside its micro-kernels
//   (the first M block writes B into SB while computing, test_temp.c:9-537,
//   test_SMM_NT.c:207-268; Kernel/SMM_m8n12_new.h:2853-3341) instead of a
//   separate PACK.c pass.  Here the same idea is the global->LDS copy: each
//   K chunk is read from HBM once with coalesced 16-byte loads along the
//   contiguous dimension and written to LDS already transposed into the
//   k-contiguous panel layout the MFMA operand reads want, so NN / NT / TN /
//   TT all share one compute loop.  For the BF16 modes the hi/lo (/lo2) split
//   is done in the same copy, so HBM traffic stays that of FP32 inputs.
// * Edge tiles.  The reference handles M/N/K remainders with separate tail
//   kernels (M 4/3/2/1 and N 8/4/2/1 in test_temp_L1.c:3998-5005, the 4x17
//   odd-N kernel in SMM_m4n17.c); here tails are predicated inside the same
//   kernel: out-of-range loads become zeros, out-of-range stores are masked,
//   16x16 MFMA sub-tiles entirely outside C are skipped, and for FP32/FP64
//   the K chunk's LDS layout is permuted so the MFMA k-steps past K are
//   skipped too.
// * Register tiling.  Each wave owns one quadrant of the tile, BM/32 x BN/32
//   sub-tiles of 16x16, multiplied without per-sub-tile branches; a quadrant
//   outside the job is skipped as a whole.
// * Pipelining.  A workgroup runs a stream of jobs (tiles of GEMMs) as one
//   flat sequence of K chunks: two chunks are in flight in registers while
//   the current one is multiplied out of LDS, across job boundaries; LDS is
//   double-buffered when it fits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ls {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

enum Prec : int { P_FP32 = 0, P_BF16X3 = 1, P_BF16 = 2, P_BF16X6 = 3, P_FP64 = 4 };

constexpr int kWave = 64;

// One GEMM.  Leading dimensions are in elements of the storage.
struct GemmDesc {
    const void *A;
    const void *B;
    void *C;
    int M, N, K;
    int lda, ldb, ldc;
    int flags;  // kChain: the next job of the workgroup's stream continues this accumulator
};

// A 2-D convolution whose im2col matrix is the B operand of a GEMM, generated
// by the operand loads instead of stored.  Input X is NCHW, one image's
// channels C x H x W; the filter is R x S; output pixel (p, q) of a P x Q
// map reads X[c][p*sh - ph + r*dh][q*sw - pw + s*dw] for filter tap
// (c, r, s), or 0 where that falls in the padding.  The im2col matrix of an
// image is (C*R*S) x (P*Q): row t = (c*R + r)*S + s, column n = p*Q + q.
//
// Fractionally strided form (th or tw > 1; the input gradient of a strided
// convolution, strides up to 8): tap (r, s) of output pixel (p, q)
// reads X[c][(p*sh - ph + r*dh) / th][(q*sw - pw + s*dw) / tw] when both
// divisions are exact, and contributes nothing otherwise.
struct ConvGeom {
    int C, H, W;            // input channels, height, width
    int R, S, RS;           // filter height, width; R x S
    int Q;                  // output width
    int sh, sw, ph, pw, dh, dw;  // stride, padding, dilation
    int th, tw;             // fractional strides (1: a plain convolution)
    unsigned mRS, mS, mQ;   // ceil(2^32 / d) for d = RS, S, Q; 0 when d == 1
    unsigned mTh, mTw;      // the same for th, tw
    // Batch folded into the pixel dimension (fold = the number of images, 0
    // when not folded): output pixel n of the GEMM is pixel n % PQ of image
    // n / PQ, whose input starts img elements after the previous image's.
    int fold, PQ, img;
    unsigned mPQ;
};

// floor(x / d) for d > 0 and any sign of x.
__device__ __forceinline__ int floordiv(int x, int d) { return x >= 0 ? x / d : -((d - 1 - x) / d); }

// How a job source's B operand is formed: read from memory; gathered as a
// convolution's im2col matrix (k = filter tap, column = output pixel); or,
// for a B stored transposed, gathered as the im2col matrix itself (stored
// row = filter tap, k = output pixel).  kIm2colTR is kIm2colT for a
// convolution of stride 1 along W whose output width Q is a multiple of 4:
// a thread's four consecutive pixels then share one output row and read
// four consecutive input columns, taken by one 16-byte load.
enum BSource : int { kPlainB = 0, kIm2col = 1, kIm2colT = 2, kIm2colTR = 3 };

// x / d for the multiplier m = ceil(2^32 / d) (m == 0: d == 1).  Exact when
// x * d < 2^32: with e = m*d - 2^32 < d, x*m / 2^32 = x/d + x*e / (d 2^32),
// and x*e < 2^32 keeps the excess below 1/d.
// Written as a select rather than a branch so that scalar uses stay
// straight-line code.
__device__ __forceinline__ int udiv(int x, unsigned m) { return (int)__umulhi((unsigned)x, m) + (m == 0u ? x : 0); }

// Where column col of a tile's C lands, relative to the tile's row origin.
// PlainCols: the tile's C pointer already includes its first column.
// Rows past M fall past the buffer's extent ((M - 1) * ldo + N elements).
struct PlainCols {
    static constexpr bool MASK_ROWS = false;
    __device__ __forceinline__ int operator()(int col) const { return col; }
    __device__ __forceinline__ int extent(int M, int ldo, int N) const { return (M - 1) * ldo + N; }
};

// The output of a convolution with the batch folded into its pixels
// (NCHW, Cout rows of PQ pixels per image): column n0 + col is pixel
// n % PQ of image n / PQ, at n + (n / PQ) * skip with skip = (Cout - 1) * PQ;
// ext is the number of elements from the row origin to the output's end,
// so the stores mask rows past M themselves (they would land on other
// images' pixels).
struct FoldedCols {
    static constexpr bool MASK_ROWS = true;
    int n0, skip, ext;
    unsigned mPQ;
    __device__ __forceinline__ int operator()(int col) const
    {
        const int n = n0 + col;
        return n + udiv(n, mPQ) * skip;
    }
    __device__ __forceinline__ int extent(int, int, int) const { return ext; }
};

// Where a split-mode stream recomputes its Inf / NaN results in FP32 (see
// Tile::any_nonfinite): at the store of the tile, leaving those elements out
// of it (the one C tile of a single-job stream, which may scale by alpha and
// beta); after the stream, reading back the stored tiles of the jobs that
// held one (an out-of-line call); or by the stream's caller, told so by
// stream()'s result.  Each source takes the one that left its kernels'
// register counts as they were.
enum RepairAt : int { kRepairAtStore = 0, kRepairAfterCall = 1, kRepairByCaller = 2 };

// GemmDesc::flags bit: the job is one segment of a longer reduction.  Its C
// tile is not stored and the next job of the same stream (which must have the
// same tile) adds to the same accumulators.
constexpr int kChain = 1;

template <int PREC> struct PrecInfo;
template <> struct PrecInfo<P_FP32> { typedef float T; static constexpr int planes = 0; };
template <> struct PrecInfo<P_FP64> { typedef double T; static constexpr int planes = 0; };
template <> struct PrecInfo<P_BF16> { typedef float T; static constexpr int planes = 1; };
template <> struct PrecInfo<P_BF16X3> { typedef float T; static constexpr int planes = 2; };
template <> struct PrecInfo<P_BF16X6> { typedef float T; static constexpr int planes = 3; };

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Global-memory (address space 1) views of the operand and result pointers.
// Through a generic pointer the compiler emits flat_load / flat_store, which
// also count in lgkmcnt, so every wait on an LDS read would drain the global
// prefetches as well; global_load / global_store keep the two apart.
template <typename T> using gptr = __attribute__((address_space(1))) T *;
template <typename T> using gcptr = const __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gcptr<T> to_global(const T *p) { return (gcptr<T>)p; }
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T *p) { return (gptr<T>)p; }

// Raw buffer resources.  Loads and stores go through buffer instructions with
// a per-lane byte offset; a lane that must not touch memory gets kOOB, an
// offset past num_records, for which the hardware returns zeros (loads) or
// drops the write (stores).  So no load or store sits in a divergent branch,
// the number of memory instructions per step is static, and the compiler can
// wait for one chunk's loads (vmcnt(n)) while the next chunk's stay in
// flight.  Offsets are 32-bit: leading dimensions are limited to < 2^21.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int kOOB = (int)0x80000000u;

// off if ok, else an offset past num_records (bit 31 set); branch-free.
__device__ __forceinline__ int oob_unless(bool ok, int off) { return off | (int)((unsigned)(!ok) << 31); }

// nrec: bytes addressable from base.  The range check is per dword (measured
// on gfx950): a 16-byte load that runs past nrec returns its in-range dwords
// and zeros for the rest, and 16-byte loads need only 4-byte alignment.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, int nrec = 0x7FFFFFFF)
{
    // stride 0, word 3 = gfx9-family data format
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, nrec, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T buf_load1(__amdgpu_buffer_rsrc_t r, int off)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// 16 bytes: four floats or two doubles.  The whole vector is bit-cast
// before its elements are taken: extracting lanes of the u32x4 result one by
// one made the compiler shrink the load to a single dword.
template <typename T>
__device__ __forceinline__ void buf_load16(__amdgpu_buffer_rsrc_t r, int off, T *v)
{
    if constexpr (sizeof(T) == 4) {
        const f32x4 x = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
        v[0] = x[0];
        v[1] = x[1];
        v[2] = x[2];
        v[3] = x[3];
    } else {
        const f64x2 x = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
        v[0] = x[0];
        v[1] = x[1];
    }
}

// Result stores carry the sc0 cache-policy bit (cache policy operand bit 0 on
// gfx950), measured ~2% faster than the default policy on the batched
// workload; nt or sc1 stores were much slower there.
template <typename T>
__device__ __forceinline__ void buf_store1(__amdgpu_buffer_rsrc_t r, int off, T x)
{
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, off, 0, 1);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, off, 0, 1);
}

// Split four floats into NP bf16 pieces each, x ~= p0 + p1 + p2, every piece
// the round-to-nearest-even of the remaining residual (the residual
// subtraction is exact in FP32).  Pieces come packed two per dword (element
// 2j in the low half), which is both the LDS layout and what
// v_cvt_pk_bf16_f32 produces; a piece is widened back to FP32 by a shift or a
// mask of its dword, so each float costs three VALU operations per piece
// after the first.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned pack_bf16(float a, float b)
{
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

template <int NP>
__device__ __forceinline__ void split4_bf16(const float (&x)[4], u32x2 (&p)[3])
{
    float r[4] = {x[0], x[1], x[2], x[3]};
#pragma unroll
    for (int pl = 0; pl < NP; ++pl) {
        const unsigned w0 = pack_bf16(r[0], r[1]), w1 = pack_bf16(r[2], r[3]);
        p[pl] = u32x2{w0, w1};
        if (pl + 1 < NP) {
            r[0] -= __builtin_bit_cast(float, w0 << 16);
            r[1] -= __builtin_bit_cast(float, w0 & 0xffff0000u);
            r[2] -= __builtin_bit_cast(float, w1 << 16);
            r[3] -= __builtin_bit_cast(float, w1 & 0xffff0000u);
        }
    }
}

// Position of k (0..31, bf16 elements) inside row `row` of a SPLIT panel
// plane: the 16-byte chunk k / 8 is swizzled with bits 1-2 of the row.
__device__ __forceinline__ int swz_k(int row, int k) { return (((k >> 3) ^ ((row >> 1) & 3)) << 3) | (k & 7); }

// ---------------------------------------------------------------------------
// LDS panel layouts of one operand chunk (ROWS x BK; "rows" are M for A, N
// for B).  Which one is used depends on the mode and on whether the operand
// is contiguous along k in memory (A NoTrans, B Trans) or along rows (A
// Trans, B NoTrans).  Each is chosen so both the store from registers and
// the MFMA operand reads stay (nearly) free of LDS bank conflicts.
//   ROWK  FP32/FP64, k-contiguous operand: [row][BK + pad], k permuted (kpos)
//         so a lane's k steps are one contiguous 16-byte run.
//   KROW  FP32/FP64, row-contiguous operand: [k][ROWS + 16], a straight copy
//         of the 16-byte global loads; the MFMA reads one element per k step.
//   SPLIT BF16 modes: NP planes of [row][BK] bf16 (64-byte rows, no
//         padding), the four 16-byte chunks of row r (k = 8c..8c+7) stored
//         in chunk order c ^ ((r >> 1) & 3); the threads hold four
//         consecutive k of one row (16-byte loads for a k-contiguous operand,
//         lane-per-row 4-byte loads otherwise) and write one 8-byte run per
//         plane.  Measured on gfx950 (SQ_LDS_BANK_CONFLICT): with this
//         swizzle the MFMA operand reads (lane (r, q): 16 bytes of row r,
//         chunk q) and the 8-lanes-per-row writes are conflict-free, and the
//         lane-per-row writes are 2-way; an 80-byte padded row made all three
//         2-way.
// ---------------------------------------------------------------------------
enum Layout : int { L_ROWK = 0, L_KROW = 1, L_SPLIT = 2 };

template <typename T, int NP, int ROWS, int BK, bool KCONTIG>
struct PanelInfo {
    static constexpr int LAYOUT = NP > 0 ? L_SPLIT : (KCONTIG ? L_ROWK : L_KROW);
    static constexpr int VEC = 16 / sizeof(T);
    static constexpr int LDK = BK + VEC;    // ROWK row length (elements)
    static constexpr int LDR = ROWS + 16;   // KROW row length (elements)
    static constexpr int LDKH = BK;         // SPLIT row length (bf16)
    static constexpr int BYTES = LAYOUT == L_SPLIT ? NP * ROWS * LDKH * 2
                                 : LAYOUT == L_ROWK ? ROWS * LDK * (int)sizeof(T)
                                                    : BK * LDR * (int)sizeof(T);
};

// ---------------------------------------------------------------------------
// Configuration of one tile kernel.
// ---------------------------------------------------------------------------
// PERSIST_: the configuration of the persistent batched kernel (its wave
// placement and LDS buffering may differ, below).
template <int PREC_, int TA_, int TB_, int BM_, int BN_, bool PERSIST_ = false>
struct TileCfg {
    static constexpr int PREC = PREC_, TA = TA_, TB = TB_, BM = BM_, BN = BN_;
    typedef typename PrecInfo<PREC>::T T;
    static constexpr int NP = PrecInfo<PREC>::planes;  // 0: no split (FP32/FP64 MFMA)
    static constexpr bool SPLIT = NP > 0;
    static constexpr int BK = sizeof(T) == 8 ? 16 : 32;
    static constexpr int VEC = 16 / sizeof(T);            // elements per 16-byte load
    static constexpr int NSLOT = BK / 4;                  // 16x16x4 MFMA k-steps per chunk
    static constexpr bool A_KCONTIG = !TA;  // A[m][k] rows contiguous along k
    static constexpr bool B_KCONTIG = TB;   // B^T[n][k] contiguous along k
    typedef PanelInfo<T, NP, BM, BK, A_KCONTIG> PA;
    typedef PanelInfo<T, NP, BN, BK, B_KCONTIG> PB;
    static constexpr int A_BYTES = PA::BYTES;
    static constexpr int B_BYTES = PB::BYTES;
    static constexpr int BUF_BYTES = A_BYTES + B_BYTES;
    // Two LDS buffers (the next chunk written while the current one is
    // multiplied) where they fit twice per CU.  BF16X6's persistent 128 x 128
    // kernel, whose three bf16 planes take 48 KB a buffer, keeps one so that
    // two workgroups share a CU: on the 50k U{1..128} batch its launch ran
    // 0.738 ms with two buffers and one workgroup per CU, 0.583 ms with one
    // buffer and two.  (The single-GEMM and slice kernels hold more than 128
    // registers, so one workgroup per CU, and keep both buffers.)
    static constexpr int NBUF = (PERSIST_ && PREC_ == P_BF16X6 && BM_ == 128 && BN_ == 128) ? 1
                                : (2 * BUF_BYTES <= 96 * 1024) ? 2 : 1;
    static constexpr int LDS_BYTES = NBUF * BUF_BYTES;
    // Workgroup size: eight waves for the 128 x 128 tiles, four otherwise.
    // The waves form a GM x GN grid over the tile, each owning WM x WN
    // sub-tiles of 16 x 16: two columns of waves for square tiles, one column
    // (the rows split) for tall tiles and one row for wide ones, so the tall-
    // and short-skinny GEMMs those tiles serve keep every wave on valid rows
    // or columns.
    // (The 16 x 16 tile, used for the smallest FP64 GEMMs of a batch, is one
    // wave.)
    static constexpr int NT = (BM == 128 && BN == 128) ? 512 : (BM == 16 && BN == 16 ? 64 : 256);
    static constexpr int NW = NT / 64;
    static constexpr int GN = BM == BN ? (NW >= 2 ? 2 : 1) : (BM > BN ? 1 : NW), GM = NW / GN;
    static constexpr int WM = BM / 16 / GM, WN = BN / 16 / GN;
    // Which sub-tiles a wave owns.  FP32 / FP64, and the split modes'
    // persistent batched kernels (INTERLEAVE): wave (g, h) of the grid owns
    // the 16-row blocks g, g + GM, ... and the 16-column blocks h, h + GN,
    // ...; a GEMM smaller than the tile covers its top-left corner, and
    // interleaved blocks spread that corner over the waves, whose MFMAs skip
    // the sub-tiles outside the GEMM (on the 50k U{1..128} batch the 128 x
    // 128 launch went from 0.593 to 0.556 ms in BF16X6 and from 0.529 to
    // 0.521 ms in BF16X3).  The split modes' single-GEMM and slice kernels,
    // whose tiles are mostly full, keep contiguous blocks (rows g*WM .. g*WM
    // + WM - 1) and multiply every sub-tile of a wave holding part of the
    // GEMM: interleaved, the BF16X3 conv launches ran 7-12% slower.  In an eight-wave workgroup waves w and w + 4 run on one SIMD
    // (the hardware's placement, read from HW_ID on gfx950), so wave w takes
    // grid place g = w / 2, h = (w & 1) ^ (w / 4): each SIMD then holds the
    // blocks (g, h) and (g + 2, 1 - h), and the corner of a GEMM splits
    // evenly over the four SIMDs (on the 50k U{1..128} batch the busiest
    // SIMD's share of the 128 x 128 tiles' MFMAs falls from 42% to 28%).
    static constexpr bool INTERLEAVE = !SPLIT || PERSIST_;
    // Memory phases of the chunk pipeline at raised wave priority (Tile::step):
    // the FP32 persistent kernels and every split-mode kernel.
    static constexpr bool PRIO = PREC != P_FP64 && (PERSIST_ || SPLIT);
    __host__ __device__ static constexpr int wave_g(int w)
    {
        return NW == 8 && INTERLEAVE ? w >> 1 : w / GN;
    }
    __host__ __device__ static constexpr int wave_h(int w)
    {
        return NW == 8 && INTERLEAVE ? (w & 1) ^ (w >> 2) : w % GN;
    }
    // 16-row (16-column) block of the wave's a-th sub-tile row (column).
    __host__ __device__ static constexpr int row_block(int g, int a) { return INTERLEAVE ? g + a * GM : g * WM + a; }
    __host__ __device__ static constexpr int col_block(int h, int b) { return INTERLEAVE ? h + b * GN : h * WN + b; }
    // How many of the wave's sub-tile rows (columns) lie in the first mt
    // (nt) blocks -- always a prefix of them.
    __host__ __device__ static constexpr int rows_in(int g, int mt)
    {
        return INTERLEAVE ? (g < mt ? ((mt - g + GM - 1) / GM < WM ? (mt - g + GM - 1) / GM : WM) : 0)
                          : (mt - g * WM <= 0 ? 0 : (mt - g * WM < WM ? mt - g * WM : WM));
    }
    __host__ __device__ static constexpr int cols_in(int h, int nt)
    {
        return INTERLEAVE ? (h < nt ? ((nt - h + GN - 1) / GN < WN ? (nt - h + GN - 1) / GN : WN) : 0)
                          : (nt - h * WN <= 0 ? 0 : (nt - h * WN < WN ? nt - h * WN : WN));
    }
    static_assert(NW != 8 || (GM == 4 && GN == 2), "the eight-wave placement assumes a 4 x 2 grid");
    // Waves per SIMD the persistent kernel is compiled for: four for the
    // eight-wave tiles of FP32 / BF16 / BF16X3 / BF16X6 (two workgroups per
    // CU, at most 128 registers), two for the other 128 x 128 tiles.  (The
    // single-GEMM tile kernel, with its alpha / beta epilogue, would spill
    // under the four-wave limit and keeps the default.)
    static constexpr int WAVES_PER_EU = (NT == 512 && PREC != P_FP64) ? 4
                                        : (BM * BN >= 128 * 128 ? 2 : 1);
    typedef typename std::conditional<sizeof(T) == 8, f64x4, f32x4>::type Acc;
    static_assert(BM % 16 == 0 && BN % 16 == 0, "tile must be a multiple of 16");
    static_assert(WM >= 1 && WN >= 1 && WM * GM * 16 == BM && WN * GN * 16 == BN, "the wave grid covers the tile");
    static_assert(WM * WN <= 16, "at most 16 sub-tiles per wave");
    static_assert(A_BYTES % 16 == 0 && B_BYTES % 16 == 0, "panels stay 16-byte aligned");
};

// LDS position of k inside a chunk for the ROWK layout: lane group q of a
// 16x16x4 MFMA takes k = 4*s + q at step s, and a lane reads its NSLOT values
// for steps 0..NSLOT-1 as one contiguous run, so k = 4*s + q is stored at
// q*NSLOT + s.  Steps whose k are all >= K can then be skipped.
template <int NSLOT>
__device__ __forceinline__ int kpos(int k) { return (k & 3) * NSLOT + (k >> 2); }

// ---------------------------------------------------------------------------
// Global -> register staging of one operand panel chunk (ROWS x BK).
// KCONTIG: storage is [row][k] (stride ld); otherwise [k][row].  Each thread
// holds groups of VEC elements consecutive along the contiguous dimension,
// loaded by one 16-byte buffer load (any 4-byte alignment); rows past the
// tile take an out-of-range offset and elements past the K / row tail are
// zeroed by a select.  The buffer's num_records is the job's exact extent,
// so a group running past the end of the matrix reads nothing beyond it.
// LANEROW (BF16 modes and the im2col operand; row-contiguous operands):
// lane-per-row 4-byte loads, each thread holding VEC consecutive k of one
// row, so the lanes of a wave share their k.
// ---------------------------------------------------------------------------
template <typename T, int NP, int ROWS, int BK, bool KCONTIG, int NT, bool LR = (NP > 0 && !KCONTIG)>
struct Stage {
    typedef PanelInfo<T, NP, ROWS, BK, KCONTIG> P;
    static constexpr int VEC = 16 / sizeof(T);
    static constexpr bool LANEROW = LR;
    static_assert(!(LR && KCONTIG), "lane-per-row loads are for row-contiguous operands");
    static constexpr int NVEC = ROWS * BK / VEC;
    static constexpr int PER = (NVEC + NT - 1) / NT;
    static constexpr int ES = sizeof(T);
    T v[PER][VEC];
    int kcut;  // valid k of the chunk (k-contiguous panels mask the rest at the LDS write)
    int run[PER];  // 16-byte im2col runs: leading padding lo | valid end hi << 8

    // First element (row, k) of the thread's group of VEC elements.
    __device__ __forceinline__ static void coords(int idx, int &row, int &k)
    {
        if constexpr (KCONTIG) {
            row = idx / (BK / VEC);
            k = (idx % (BK / VEC)) * VEC;
        } else if constexpr (LANEROW) {
            row = idx % ROWS;
            k = (idx / ROWS) * VEC;
        } else {
            k = idx / (ROWS / VEC);
            row = (idx % (ROWS / VEC)) * VEC;
        }
    }

    // Load the chunk whose first k is k0 (rows [0, R), k < kend of the job).
    // live == false issues the same loads with nothing in range.  The buffer
    // covers exactly the job's part of the chunk (through row R-1 and k
    // kc-1), and ld >= R, ld >= kc, so most out-of-tile elements fall past
    // num_records on their own: rows >= R of a k-contiguous panel, k >= kc
    // of a row-contiguous one.  Only what is in range but outside the tile
    // (k >= kc inside a row, rows >= R inside a k row) is masked.
    __device__ __forceinline__ void load(const T *__restrict__ base, int ld, int R, int kend, int k0, bool live,
                                         int tid)
    {
        const int kc = min(kend - k0, BK);  // valid k in this chunk
        kcut = kc;
        const int nrec = (live && kc > 0 && R > 0) ? (KCONTIG ? ((R - 1) * ld + kc) * ES : ((kc - 1) * ld + R) * ES)
                                                   : 0;
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(KCONTIG ? base + k0 : base + (size_t)k0 * ld, nrec);
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = tid + i * NT;
            int row, k;
            coords(idx, row, k);
            const bool act = (NVEC % NT == 0) | (idx < NVEC);
            if constexpr (KCONTIG) {
                buf_load16<T>(rs, oob_unless(act & (k < kc), (row * ld + k) * ES), v[i]);
            } else if constexpr (LANEROW) {
                const int hi = (act & (row < R)) ? 0 : kOOB;
                const int off = (k * ld + row) * ES;
#pragma unroll
                for (int e = 0; e < VEC; ++e) v[i][e] = buf_load1<T>(rs, (off + e * ld * ES) | hi);
            } else {
                buf_load16<T>(rs, oob_unless(act, (k * ld + row) * ES), v[i]);
#pragma unroll
                for (int e = 0; e < VEC; ++e) v[i][e] = (row + e < R) ? v[i][e] : T(0);
            }
        }
    }

    // Per-lane coordinates of the im2col rows (output pixels) this thread
    // loads, fixed for one output tile: the input position (ih0, iw0) of
    // tap (0, 0), its byte offset in channel 0, and -- for filters of at
    // most 32 taps -- the mask of taps (bit r*S + s) that fall inside the
    // image.  A row past the tile (or the panel) gets ih0 = 2^29, which no
    // tap brings back into the image, and an empty mask.  A fractionally
    // strided gather with more than 32 taps keeps the pixel's remainders
    // al | al2 << 16 (below) in place of the mask.
    struct Pixels {
        int ih0[PER], iw0[PER], base[PER];
        unsigned taps[PER];
    };

    __device__ __forceinline__ static Pixels pixels(const ConvGeom &g, int n0, int N, int tid)
    {
        Pixels px;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = tid + i * NT;
            int row, k;
            coords(idx, row, k);
            const bool act = (NVEC % NT == 0) | (idx < NVEC);
            int n = n0 + row, img_off = 0;
            if (g.fold) {  // image n / PQ of the batch
                const int b = udiv(n, g.mPQ);
                n -= b * g.PQ;
                img_off = b * g.img;
            }
            const int p = udiv(n, g.mQ), q = n - p * g.Q;
            const bool rok = act & (row < N);
            int ih0, iw0;
            unsigned taps = 0;
            if (g.th == 1 && g.tw == 1) {
                ih0 = rok ? p * g.sh - g.ph : (1 << 29);
                iw0 = q * g.sw - g.pw;
                if (g.RS <= 32) {
                    unsigned cols = 0;
                    for (int s = 0; s < g.S; ++s)
                        cols |= (unsigned)((unsigned)(iw0 + s * g.dw) < (unsigned)g.W) << s;
                    for (int r = 0; r < g.R; ++r)
                        if ((unsigned)(ih0 + r * g.dh) < (unsigned)g.H) taps |= cols << (r * g.S);
                }
            } else {
                // u = p*sh - ph = a*th + al (floor division); a tap with
                // r*dh = b*th + be contributes when al + be is 0 or th, and
                // then reads input row a + (al > 0) + b: the row splits into
                // a per-pixel part (kept in ih0) and a per-tap part b.
                const int u = p * g.sh - g.ph, v = q * g.sw - g.pw;
                const int a = floordiv(u, g.th), al = u - a * g.th;
                const int a2 = floordiv(v, g.tw), al2 = v - a2 * g.tw;
                ih0 = rok ? a + (al > 0) : (1 << 29);
                iw0 = a2 + (al2 > 0);
                if (g.RS <= 32) {
                    unsigned cols = 0;
                    for (int s = 0; s < g.S; ++s) {
                        const int b = s * g.dw / g.tw, be = s * g.dw - b * g.tw;
                        const bool hit = (al2 + be == 0) | (al2 + be == g.tw);
                        cols |= (unsigned)(hit & ((unsigned)(iw0 + b) < (unsigned)g.W)) << s;
                    }
                    for (int r = 0; r < g.R; ++r) {
                        const int b = r * g.dh / g.th, be = r * g.dh - b * g.th;
                        const bool hit = (al + be == 0) | (al + be == g.th);
                        if (hit & ((unsigned)(ih0 + b) < (unsigned)g.H)) taps |= cols << (r * g.S);
                    }
                } else {
                    taps = (unsigned)al | ((unsigned)al2 << 16);
                }
            }
            // modulo 2^32: only used where ih0 is a real row; kept opaque so
            // that the compiler adds a tap's offset to it instead of
            // re-deriving (ih0 + dy) * W with a quarter-rate multiply
            int base = (int)(((unsigned)img_off + (unsigned)ih0 * (unsigned)g.W + (unsigned)iw0) * ES);
            asm volatile("" : "+v"(base));
            px.ih0[i] = ih0;
            px.iw0[i] = iw0;
            px.base[i] = base;
            px.taps[i] = taps;
        }
        return px;
    }

    // The chunk at k0 of a convolution's im2col panel (a row-contiguous
    // operand: row = output pixel, k = filter tap) for the tile whose lanes
    // are px.  X points at the job's first channel and kend is the job's
    // taps.  Every element is one 4-byte load; taps in the padding, rows
    // past the tile and taps past kend take an out-of-range offset and read
    // 0.  The lanes of a wave hold one group of VEC taps (lane-per-row
    // layout, ROWS a multiple of the wave).  Lane l decodes tap k0 + l % BK
    // of the chunk (channel, filter row and column: one division each, all
    // lanes at once) and the wave reads the decoded taps it holds from those
    // lanes (v_readlane), so per element a filter of at most 32 taps costs a
    // test of the tap's bit in the pixel's mask, a larger one two adds and
    // two range compares (and, fractionally strided, the test that both
    // remainders sum to 0 or the stride).  (Decoding in scalar registers made the CU's one
    // scalar unit the bottleneck.)  The offsets are formed first and the
    // loads issued after, outside the branch on the filter size.
    __device__ __forceinline__ void load_im2col(const T *__restrict__ X, const ConvGeom &g, const Pixels &px,
                                                int kend, int k0, bool live, int tid)
    {
        static_assert(LANEROW && ROWS % kWave == 0, "a wave's lanes share their taps");
        static_assert(BK <= kWave, "one lane per tap of the chunk");
        const int kc = min(kend - k0, BK);
        kcut = kc;
        const int HW = g.H * g.W;
        const int nch = udiv(kend + g.RS - 1, g.mRS);  // channels the job spans
        // (folded: the job's channels of every image up to the last)
        const int span = g.fold ? (g.fold - 1) * g.img + nch * HW : nch * HW;
        const __amdgpu_buffer_rsrc_t rx = make_rsrc(X, (live && kc > 0) ? span * ES : 0);
        const int tl = k0 + (tid & (BK - 1));
        const int tc = udiv(tl, g.mRS), trem = tl - tc * g.RS;
        const int tr = udiv(trem, g.mS), ts = trem - tr * g.S;
        // the tap's input row / column offsets (fractionally strided: the
        // per-tap parts r*dh / th, s*dw / tw)
        const int tdy = udiv(tr * g.dh, g.mTh), tdx = udiv(ts * g.dw, g.mTw);
        const int ltoff = (tc * HW + tdy * g.W + tdx) * ES;
        int off[PER][VEC];
        if (g.RS <= 32) {
            const unsigned lbit = tl < kend ? 1u << trem : 0u;  // no bit for a tap past K
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                int row, k;
                coords(tid + i * NT, row, k);
                const int kw = uniform(k);
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    const unsigned bit = __builtin_amdgcn_readlane(lbit, kw + e);
                    const int toff = __builtin_amdgcn_readlane(ltoff, kw + e);
                    off[i][e] = (px.taps[i] & bit) ? px.base[i] + toff : kOOB;
                }
            }
        } else if (g.th != 1 || g.tw != 1) {
            // fractionally strided, more than 32 taps: the tap's row and
            // column parts b (tdy, tdx) and remainders be, the pixel's
            // remainders al in px.taps (al | al2 << 16, each below the
            // stride, so the packed sum keeps the halves apart); the tap
            // contributes when al + be is 0 or the stride in both
            // directions.  The row sentinel is as below.
            const int ldy = tl < kend ? tdy : (1 << 29), ldx = tdx;
            const int lbe = (tr * g.dh - tdy * g.th) | ((ts * g.dw - tdx * g.tw) << 16);
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                int row, k;
                coords(tid + i * NT, row, k);
                const int kw = uniform(k);
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    const int dy = __builtin_amdgcn_readlane(ldy, kw + e), dx = __builtin_amdgcn_readlane(ldx, kw + e);
                    const int toff = __builtin_amdgcn_readlane(ltoff, kw + e);
                    const int sum = (int)px.taps[i] + __builtin_amdgcn_readlane(lbe, kw + e);
                    const int sy = sum & 0xffff, sx = sum >> 16;
                    const bool ok = ((unsigned)(px.ih0[i] + dy) < (unsigned)g.H) &
                                    ((unsigned)(px.iw0[i] + dx) < (unsigned)g.W) & ((sy == 0) | (sy == g.th)) &
                                    ((sx == 0) | (sx == g.tw));
                    off[i][e] = ok ? px.base[i] + toff : kOOB;
                }
            }
        } else {
            // a tap past K gets the out-of-range row offset 2^29: H + 2 ph
            // < 2^29 (checked on the host) keeps every real ih0 and dy below
            // 2^29 in magnitude, so a sum with a sentinel lies in
            // [2^29 - ph, 2^30] -- above H, without overflow
            const int ldy = tl < kend ? tr * g.dh : (1 << 29), ldx = ts * g.dw;
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                int row, k;
                coords(tid + i * NT, row, k);
                const int kw = uniform(k);
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    const int dy = __builtin_amdgcn_readlane(ldy, kw + e), dx = __builtin_amdgcn_readlane(ldx, kw + e);
                    const int toff = __builtin_amdgcn_readlane(ltoff, kw + e);
                    const bool ok = ((unsigned)(px.ih0[i] + dy) < (unsigned)g.H) &
                                    ((unsigned)(px.iw0[i] + dx) < (unsigned)g.W);
                    off[i][e] = ok ? px.base[i] + toff : kOOB;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < PER; ++i)
#pragma unroll
            for (int e = 0; e < VEC; ++e) v[i][e] = buf_load1<T>(rx, off[i][e]);
    }

    // Per-lane filter taps of the rows this thread loads from a transposed
    // im2col panel (row = tap t0 + row, k = output pixel), fixed for one
    // output tile: the tap's byte offset (c*H*W + r*dh*W + s*dw) * 4 and its
    // row / column offsets r*dh, s*dw.  A row past the tile (or the panel)
    // gets dy = 2^29, which keeps every pixel out of the image.
    struct Taps {
        int toff[PER], dy[PER], dx[PER];
    };

    __device__ __forceinline__ static Taps taps(const ConvGeom &g, int t0, int N, int tid)
    {
        static_assert(KCONTIG && !LANEROW, "the transposed im2col panel is k-contiguous");
        Taps tp;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = tid + i * NT;
            int row, k;
            coords(idx, row, k);
            const bool act = (NVEC % NT == 0) | (idx < NVEC);
            const int t = t0 + row;
            const int c = udiv(t, g.mRS), rem = t - c * g.RS;
            const int r = udiv(rem, g.mS), s = rem - r * g.S;
            tp.dy[i] = (act & (row < N)) ? r * g.dh : (1 << 29);
            tp.dx[i] = s * g.dw;
            tp.toff[i] = (int)(((unsigned)c * (unsigned)(g.H * g.W) + (unsigned)(r * g.dh * g.W + s * g.dw)) * ES);
        }
        return tp;
    }

    // The chunk at k0 of a transposed im2col panel: element (row, k) is
    // X[c][p*sh - ph + r*dh][q*sw - pw + s*dw] for the row's tap (c, r, s)
    // and output pixel n0 + k0 + k = p*Q + q, or 0 in the padding and past
    // kend pixels.  Each thread holds VEC consecutive pixels of one tap:
    // the first is decoded (one division), the others follow by stepping
    // along the output row.
    //
    // RUNS (stride 1 along W, Q and n0 multiples of VEC): the VEC pixels
    // share one output row, so the thread loads the VEC input columns of the
    // tap with one 16-byte load that starts at the first one inside the
    // image (never before the channel's start), and fix_runs() moves them
    // into place once they have landed: lo leading columns in the left
    // padding are zero, and so is everything from hi on (right padding or
    // the chunk's end).  The load's columns past the row belong to the next
    // row or channel and are among those zeroed.
    template <bool RUNS = false>
    __device__ __forceinline__ void load_im2col_t(const T *__restrict__ X, const ConvGeom &g, const Taps &tp, int n0,
                                                  int kend, int k0, bool live, int tid)
    {
        static_assert(KCONTIG && !LANEROW, "the transposed im2col panel is k-contiguous");
        const int kc = min(kend - k0, BK);
        kcut = kc;
        const int HW = g.H * g.W;
        const __amdgpu_buffer_rsrc_t rx = make_rsrc(X, (live && kc > 0) ? g.C * HW * ES : 0);
        if constexpr (RUNS) {
            static_assert(VEC == 4, "runs of four 4-byte pixels");
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                int row, k;
                coords(tid + i * NT, row, k);
                const int n = n0 + k0 + k;
                const int p = udiv(n, g.mQ), q = n - p * g.Q;
                const int ih = p * g.sh - g.ph, iw = q - g.pw;
                const int iwt = iw + tp.dx[i];  // the tap's first input column
                const int lo = min(max(-iwt, 0), VEC), hi = max(min(min(g.W - iwt, kc - k), VEC), 0);
                const bool ok = ((unsigned)(ih + tp.dy[i]) < (unsigned)g.H) & (lo < hi);
                run[i] = lo | (hi << 8);
                // byte offset of column iwt + lo >= 0 of the tap's row
                const unsigned o = (unsigned)tp.toff[i] + ((unsigned)ih * (unsigned)g.W + (unsigned)(iw + lo)) * ES;
                buf_load16<T>(rx, ok ? (int)o : kOOB, v[i]);
            }
            return;
        }
        const unsigned step_col = g.sw * ES, step_row = (unsigned)(g.sh * g.W - (g.Q - 1) * g.sw) * ES;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int row, k;
            coords(tid + i * NT, row, k);
            const int n = n0 + k0 + k;
            int p = udiv(n, g.mQ), q = n - p * g.Q;
            int ih = p * g.sh - g.ph, iw = q * g.sw - g.pw;
            // byte offset of the pixel's tap (0, 0) in channel 0, modulo 2^32
            unsigned pix = ((unsigned)ih * (unsigned)g.W + (unsigned)iw) * ES;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const bool ok = (k + e < kc) & ((unsigned)(ih + tp.dy[i]) < (unsigned)g.H) &
                                ((unsigned)(iw + tp.dx[i]) < (unsigned)g.W);
                v[i][e] = buf_load1<T>(rx, ok ? (int)((unsigned)tp.toff[i] + pix) : kOOB);
                // next output pixel
                const bool wrap = q + 1 == g.Q;
                q = wrap ? 0 : q + 1;
                ih += wrap ? g.sh : 0;
                iw = wrap ? -g.pw : iw + g.sw;
                pix += wrap ? step_row : step_col;
            }
        }
    }

    // Places the columns of 16-byte im2col runs once the loads have landed:
    // element e is loaded column e - lo for lo <= e < hi, else zero.
    __device__ __forceinline__ void fix_runs()
    {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int lo = run[i] & 255, hi = run[i] >> 8;
            T w[VEC];
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                T x = v[i][e];
#pragma unroll
                for (int s = 1; s <= e; ++s) x = lo == s ? v[i][e - s] : x;
                w[e] = ((e >= lo) & (e < hi)) ? x : T(0);
            }
#pragma unroll
            for (int e = 0; e < VEC; ++e) v[i][e] = w[e];
        }
    }

    // Registers -> LDS panel.  In a k-contiguous panel the elements of a
    // group at k >= kcut (in range: the next columns of the row) are zeroed
    // here, after the loads have landed, and only in a chunk that ends
    // inside the tile's K (a uniform branch).
    __device__ __forceinline__ void store(char *__restrict__ dst, int tid)
    {
        constexpr int NSLOT = BK / 4;
        if constexpr (KCONTIG) {
            if (kcut < BK) {
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    int row, k;
                    coords(tid + i * NT, row, k);
#pragma unroll
                    for (int e = 1; e < VEC; ++e) v[i][e] = (k + e < kcut) ? v[i][e] : T(0);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = tid + i * NT;
            if (!((NVEC % NT == 0) || idx < NVEC)) continue;
            int row, k;
            coords(idx, row, k);
            if constexpr (P::LAYOUT == L_ROWK) {
                T *lds = (T *)dst;
#pragma unroll
                for (int e = 0; e < VEC; ++e) lds[row * P::LDK + kpos<NSLOT>(k + e)] = v[i][e];
            } else if constexpr (P::LAYOUT == L_KROW && LANEROW) {
                // VEC consecutive k of one row: one 4-byte write per k, the
                // wave's lanes covering 64 consecutive rows of each k row
                T *lds = (T *)dst + k * P::LDR + row;
#pragma unroll
                for (int e = 0; e < VEC; ++e) lds[e * P::LDR] = v[i][e];
            } else if constexpr (P::LAYOUT == L_KROW) {
                T *lds = (T *)dst + k * P::LDR + row;
                if constexpr (sizeof(T) == 4)
                    *(f32x4 *)lds = f32x4{v[i][0], v[i][1], v[i][2], v[i][3]};
                else
                    *(f64x2 *)lds = f64x2{v[i][0], v[i][1]};
            } else {
                static_assert(sizeof(T) == 4, "split modes take FP32 inputs");
                __bf16 *lds = (__bf16 *)dst;
                constexpr int PLANE = ROWS * P::LDKH;
                u32x2 p[3];
                split4_bf16<NP>(v[i], p);
                // four consecutive k of one row: one 8-byte write per plane
#pragma unroll
                for (int pl = 0; pl < NP; ++pl) *(u32x2 *)(lds + pl * PLANE + row * P::LDKH + swz_k(row, k)) = p[pl];
            }
        }
    }
};

// ---------------------------------------------------------------------------
// The tile computation.
// ---------------------------------------------------------------------------
template <class Cfg>
struct Tile {
    typedef typename Cfg::T T;
    typedef typename Cfg::Acc Acc;
    static constexpr int BM = Cfg::BM, BN = Cfg::BN, BK = Cfg::BK;
    static constexpr int WM = Cfg::WM, WN = Cfg::WN;
    typedef Stage<T, Cfg::NP, BM, BK, Cfg::A_KCONTIG, Cfg::NT> StageA;
    typedef Stage<T, Cfg::NP, BN, BK, Cfg::B_KCONTIG, Cfg::NT> StageB;
    // B stage of an im2col source (lane-per-row loads in every mode); the
    // transposed im2col operand uses the ordinary k-contiguous stage
    typedef Stage<T, Cfg::NP, BN, BK, false, Cfg::NT, true> StageBI;
    template <class Src>
    using StageBFor = typename std::conditional<Src::IM2COL == kIm2col, StageBI, StageB>::type;

    struct NoLanes {};
    // The per-lane coordinates a computed B operand keeps for the stream's
    // output tile (the same for every job of such a stream).
    template <class Src>
    __device__ __forceinline__ static auto lanes(const Src &src, const GemmDesc &d0, int tid)
    {
        if constexpr (Src::IM2COL == kIm2col)
            return StageBI::pixels(src.geom, d0.ldb, d0.N, tid);
        else if constexpr (Src::IM2COL == kIm2colT || Src::IM2COL == kIm2colTR)
            return StageB::taps(src.geom, src.c0, d0.N, tid);
        else
            return NoLanes{};
    }

    Acc acc[WM][WN];
    // Result scaling: with `scale` set, stores write alpha * product + beta *
    // (the stored value), reading it only when beta is nonzero.  Only the
    // single-GEMM tile kernel sets it; elsewhere it stays a constant false.
    bool scale = false;
    T alpha = T(1), beta = T(0);
    int wr0, wc0;  // this wave's place (g, h) in the wave grid (Cfg::wave_g / wave_h)
    int wm, wn;    // sub-tile rows / columns of the wave that reach into the job (a prefix)
    bool live;     // the wave holds part of the job's C tile

    // The waves own WM x WN sub-tiles each (Cfg::row_block / col_block).  In
    // the split modes every sub-tile of a live wave is multiplied and stored
    // whether or not it reaches into the job: rows and columns past the job
    // are zeros in LDS, and their stores are masked by the buffer range.
    // Issuing the MFMAs without per-sub-tile branches lets the LDS reads run
    // ahead of them; a wave wholly outside the job skips its work with one
    // uniform branch.  The FP32 and FP64 products, whose MFMAs cost 4-8x
    // more per element than the BF16 ones, skip the sub-tiles outside the
    // job (uniform branches).
    __device__ __forceinline__ void layout(int mt, int nt, int wave)
    {
        wr0 = Cfg::wave_g(wave);
        wc0 = Cfg::wave_h(wave);
        wm = Cfg::rows_in(wr0, mt);
        wn = Cfg::cols_in(wc0, nt);
        live = uniform(wm > 0 && wn > 0) != 0;
    }

    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
            for (int b = 0; b < WN; ++b) acc[a][b] = Acc{0, 0, 0, 0};
    }

    // Multiply one LDS chunk.  ksteps = number of valid 4-wide k steps (unsplit).
    // The wave's A fragments stay in registers; B fragments are read one
    // 16-column block at a time, and the MFMAs of one k step (or one split
    // product) are issued across the independent row blocks so consecutive
    // MFMAs do not wait on each other.
    __device__ __forceinline__ void compute(const char *__restrict__ buf, int lane, int ksteps)
    {
        const int r = lane & 15, q = lane >> 4;
        if (!live) return;
        if constexpr (!Cfg::SPLIT) {
            // FP32 / FP64: the MFMAs of the sub-tiles that reach into the
            // job, with uniform branches per 16-column block and per row
            // block rather than per MFMA.  Every k step of the chunk is
            // multiplied: past K the panels hold zeros, and testing each
            // step (a branch per MFMA) measured slower than the extra MFMAs
            // of a job's last chunk; a second, tested copy of the loop for
            // the last chunk did not fit the persistent kernel's registers.
            (void)ksteps;
            compute_float(buf, r, q);
        } else {
            constexpr int NP = Cfg::NP, LDKH = Cfg::PA::LDKH;
            static_assert(Cfg::PB::LDKH == LDKH, "one split row length");
            const __bf16 *As = (const __bf16 *)buf;
            const __bf16 *Bs = (const __bf16 *)(buf + Cfg::A_BYTES);
            bf16x8 af[WM][NP];
#pragma unroll
            for (int a = 0; a < WM; ++a)
#pragma unroll
                for (int pl = 0; pl < NP; ++pl)
                    af[a][pl] = *(const bf16x8 *)(As + pl * BM * LDKH + (Cfg::row_block(wr0, a) * 16 + r) * LDKH +
                                                  swz_k(r, q * 8));
            if constexpr (Cfg::INTERLEAVE) __builtin_assume(wm >= 1 && wm <= WM && wn >= 1 && wn <= WN);
            const bool all_rows = !Cfg::INTERLEAVE || wm == WM;
            // partial products (A piece, B piece), smallest first.  Two
            // pieces: lo*hi, hi*lo, hi*hi.  Three pieces (exact for every
            // FP32 value): all but lo2*lo2 (~2^-32 |a||b|) -- the two
            // products of size ~2^-24 |a||b| are needed for the FP32 bound
            // at K = 1, where one rounding already costs up to 2^-24.
            constexpr int NPROD = NP == 1 ? 1 : (NP == 2 ? 3 : 8);
            constexpr int PA[8] = {NP == 3 ? 2 : 1, NP == 3 ? 1 : 0, NP == 3 ? 2 : 0, 1, 0, 1, 0, 0};
            constexpr int PB[8] = {NP == 3 ? 1 : 0, NP == 3 ? 2 : 1, 0, 1, 2, 0, 1, 0};
#pragma unroll
            for (int b = 0; b < WN; ++b) {
                if (Cfg::INTERLEAVE && b >= wn) break;
                bf16x8 bfr[NP];
#pragma unroll
                for (int pl = 0; pl < NP; ++pl)
                    bfr[pl] = *(const bf16x8 *)(Bs + pl * BN * LDKH + (Cfg::col_block(wc0, b) * 16 + r) * LDKH +
                                                 swz_k(r, q * 8));
                if (all_rows) {
#pragma unroll
                    for (int p = 0; p < NPROD; ++p) {
                        const int ia = NP == 1 ? 0 : PA[p], ib = NP == 1 ? 0 : PB[p];
#pragma unroll
                        for (int a = 0; a < WM; ++a)
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][ia], bfr[ib], acc[a][b], 0, 0, 0);
                    }
                } else {
#pragma unroll
                    for (int p = 0; p < NPROD; ++p) {
                        const int ia = NP == 1 ? 0 : PA[p], ib = NP == 1 ? 0 : PB[p];
#pragma unroll
                        for (int a = 0; a < WM; ++a) {
                            if (a >= wm) break;
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][ia], bfr[ib], acc[a][b], 0, 0, 0);
                        }
                    }
                }
            }
        }
    }

    __device__ __forceinline__ void compute_float(const char *__restrict__ buf, int r, int q)
    {
        constexpr int NS = Cfg::NSLOT;
        const T *As = (const T *)buf;
        const T *Bs = (const T *)(buf + Cfg::A_BYTES);
        T af[WM][NS];
#pragma unroll
        for (int a = 0; a < WM; ++a) read_frag<typename Cfg::PA>(As, Cfg::row_block(wr0, a), r, q, af[a]);
        // a live wave holds at least one sub-tile row and column (telling
        // the compiler so keeps the partial-row path to the rows that can
        // be missing)
        __builtin_assume(wm >= 1 && wm <= WM && wn >= 1 && wn <= WN);
        const bool all_rows = wm == WM;
#pragma unroll
        for (int b = 0; b < WN; ++b) {
            if (b >= wn) break;
            T bf[NS];
            read_frag<typename Cfg::PB>(Bs, Cfg::col_block(wc0, b), r, q, bf);
            if (all_rows) {
#pragma unroll
                for (int s = 0; s < NS; ++s)
#pragma unroll
                    for (int a = 0; a < WM; ++a) mfma_step(acc[a][b], af[a][s], bf[s]);
            } else {
#pragma unroll
                for (int s = 0; s < NS; ++s)
#pragma unroll
                    for (int a = 0; a < WM; ++a) {
                        if (a >= wm) break;
                        mfma_step(acc[a][b], af[a][s], bf[s]);
                    }
            }
        }
    }

    __device__ __forceinline__ static void mfma_step(Acc &c, T a, T b)
    {
        if constexpr (sizeof(T) == 4)
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
        else
            c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }

    template <class StageT>
    __device__ __forceinline__ static void store_stage(StageT &st, char *dst, int tid)
    {
        st.store(dst, tid);
    }

    // The NS k-step values of one MFMA operand fragment: sub-tile row block
    // `blk` (16 rows) of a ROWK or KROW panel, for lane (r, q).
    template <class P>
    __device__ __forceinline__ static void read_frag(const T *__restrict__ X, int blk, int r, int q,
                                                     T (&f)[Cfg::NSLOT])
    {
        constexpr int NS = Cfg::NSLOT;
        if constexpr (P::LAYOUT == L_ROWK) {
            constexpr int V = 16 / sizeof(T);
            const T *p = X + (blk * 16 + r) * P::LDK + q * NS;
#pragma unroll
            for (int s = 0; s < NS; s += V) {
                if constexpr (sizeof(T) == 4) {
                    f32x4 t = *(const f32x4 *)(p + s);
                    f[s] = t[0]; f[s + 1] = t[1]; f[s + 2] = t[2]; f[s + 3] = t[3];
                } else {
                    f64x2 t = *(const f64x2 *)(p + s);
                    f[s] = t[0]; f[s + 1] = t[1];
                }
            }
        } else {
            const T *p = X + q * P::LDR + blk * 16 + r;
#pragma unroll
            for (int s = 0; s < NS; ++s) f[s] = p[4 * s * P::LDR];
        }
    }

    // Wave layout for a tile of d at (m0, n0), accumulators cleared.
    __device__ __forceinline__ void begin(int M, int N, int m0, int n0, int wave)
    {
        layout((min(BM, M - m0) + 15) >> 4, (min(BN, N - n0) + 15) >> 4, wave);
        zero();
    }

    // Global -> registers: the K chunk at k0 of job d (live == false: a
    // placeholder load that returns zeros, keeping instruction counts static).
    // A source with IM2COL == kIm2col describes B as a convolution's im2col
    // matrix: d.B is the image at the job's first channel and d.ldb the
    // tile's first output pixel.  With kIm2colT, B is stored transposed and
    // is the im2col matrix itself (rows = taps from src.c0, k = output
    // pixels from src.rho0); d.B is the image.  Every job of such a stream
    // has the same output tile, whose per-lane coordinates px are computed
    // once.
    template <class Src, class SB, class PX>
    __device__ __forceinline__ static void load_chunk(StageA &sa, SB &sb, const Src &src, const PX &px,
                                                      const GemmDesc &d, int k0, bool live, int tid)
    {
        sa.load((const T *)d.A, d.lda, d.M, d.K, k0, live, tid);
        if constexpr (Src::IM2COL == kIm2col)
            sb.load_im2col((const T *)d.B, src.geom, px, d.K, k0, live, tid);
        else if constexpr (Src::IM2COL == kIm2colT)
            sb.load_im2col_t((const T *)d.B, src.geom, px, src.rho0, d.K, k0, live, tid);
        else if constexpr (Src::IM2COL == kIm2colTR)
            sb.template load_im2col_t<true>((const T *)d.B, src.geom, px, src.rho0, d.K, k0, live, tid);
        else
            sb.load((const T *)d.B, d.ldb, d.N, d.K, k0, live, tid);
    }

    // Registers -> LDS in the packed MFMA operand layout (transpose / split).
    template <class SB>
    __device__ __forceinline__ static void chunk_to_lds(StageA &sa, SB &sb, char *buf, int tid)
    {
        store_stage(sa, buf, tid);
        store_stage(sb, buf + Cfg::A_BYTES, tid);
    }

    __device__ __forceinline__ static int ksteps(int kend, int k0) { return min(Cfg::NSLOT, (kend - k0 + 3) >> 2); }

    __device__ __forceinline__ static char *lds_buf(char *lds, int i)
    {
        return lds + (Cfg::NBUF == 2 ? (i & 1) * Cfg::BUF_BYTES : 0);
    }

    __device__ __forceinline__ static int nchunks(int K) { return max(1, (K + BK - 1) / BK); }

    // Position in the flat stream of K chunks of a workgroup's jobs.  Only
    // indices are carried; a job's descriptor is fetched from the source
    // where it is needed (scalar loads from the job array, or a few scalar
    // operations for computed sources), which keeps the scalar registers of
    // the loop few enough that uniform values are not moved to vector
    // registers.
    struct Cursor {
        int job;  // job index, -1 past the end
        int c;    // chunk within the job
        int nch;  // chunks of the job
    };

    template <class Src>
    __device__ __forceinline__ static void advance(Cursor &x, const Src &src, int stride)
    {
        if (x.job < 0) return;
        if (x.c + 1 < x.nch) {
            ++x.c;
            return;
        }
        const int nj = x.job + stride;
        if (nj < src.count()) {
            x.job = nj;
            x.c = 0;
            x.nch = nchunks(src.get(nj).K);
        } else {
            x.job = -1;
        }
    }

    // One step of the chunk pipeline.  `cur` is in LDS buffer bi; its
    // successor n1 is in flight into stage (a1, b1); stage (a0, b0) is free and
    // receives n2.  Two chunks stay in flight while `cur` is multiplied.
    template <class Src, class SB, class PX>
    __device__ __forceinline__ bool step(Cursor &cur, Cursor &n1, Cursor &n2, StageA &a0, SB &b0, StageA &a1,
                                         SB &b1, const PX &px, char *lds, int bi, const Src &src,
                                         int stride, int tid,
                                         int lane, int wave, int &bad_lo, int &bad_hi)
    {
        // Wave priorities (Cfg::PRIO): a wave issuing its loads goes first,
        // then one writing LDS or storing C, then the multiplies, so that
        // memory requests leave the CU as early as the other waves'
        // instruction streams allow (on the 50k U{1..128} batch the 128 x
        // 128 launch ran 2-3% faster in FP32 and BF16X6, and the BF16X3
        // forward of the 512 x 49 x 4608 conv layer 7%; the FP32 conv and
        // FP64 launches ran 1-2% slower and keep equal priorities).
        if constexpr (Cfg::PRIO) {
            __builtin_amdgcn_s_setprio(2);
            __builtin_amdgcn_sched_barrier(0);
        }
        load_chunk(a0, b0, src, px, src.get(max(n2.job, 0)), n2.c * BK, n2.job >= 0, tid);
        // keep the loads ahead of the multiply in the schedule
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (Cfg::PRIO) {
            __builtin_amdgcn_s_setprio(0);
            __builtin_amdgcn_sched_barrier(0);
        }
        const GemmDesc dc = src.get(cur.job);
        compute(lds_buf(lds, bi), lane, ksteps(dc.K, cur.c * BK));
        const bool chained = (dc.flags & kChain) != 0;
        // n1's chunk goes to LDS even past the end of the stream (it is then
        // the zeros of a placeholder load), so that every path has waited for
        // its loads before the C tile below is written: vector stores and
        // loads share one counter, and stores issued before that wait would
        // make it wait for n2's loads as well and drain the prefetch at every
        // job end.
        if constexpr (Cfg::NBUF == 1) __syncthreads();
        if constexpr (Src::IM2COL == kIm2colTR) b1.fix_runs();
        if constexpr (Cfg::PRIO) __builtin_amdgcn_s_setprio(1);
        chunk_to_lds(a1, b1, lds_buf(lds, bi + 1), tid);
        if (cur.c + 1 == cur.nch && !chained) {
            if constexpr (Cfg::SPLIT && Src::REPAIR == kRepairAtStore) {
                // the stream's one C tile: its Inf / NaN elements are left
                // out of the store (keeping C for beta) and recomputed right
                // away
                if (__builtin_expect(any_nonfinite(), 0)) {
                    const unsigned long long bad = nonfinite_mask();
                    store((T *)dc.C, dc.ldc, dc.M, dc.N, src.cols(), bad);
                    repair(src, cur.job, stride, bad, src.cols());
                } else {
                    store((T *)dc.C, dc.ldc, dc.M, dc.N, src.cols());
                }
            } else {
                // a job stream (no alpha / beta): the job's number is noted
                // and its Inf / NaN elements recomputed after the stream
                if constexpr (Cfg::SPLIT) {
                    if (__builtin_expect(any_nonfinite(), 0)) {
                        bad_lo = min(bad_lo, cur.job);
                        bad_hi = cur.job;
                    }
                }
                store((T *)dc.C, dc.ldc, dc.M, dc.N, src.cols());
            }
        }
        if constexpr (Cfg::PRIO) __builtin_amdgcn_s_setprio(0);
        if (n1.job < 0) return false;
        // n1 starting a job means cur ends one: a chained job keeps its sums
        if (n1.c == 0 && !chained) {
            const GemmDesc d1 = src.get(n1.job);
            begin(d1.M, d1.N, 0, 0, wave);
        }
        __syncthreads();
        cur = n1;
        n1 = n2;
        advance(n2, src, stride);
        return true;
    }

    // Job stream: this workgroup takes jobs first, first + stride, ... of src.
    // Each job is a GEMM of at most BM x BN whose pointers are already offset
    // to its tile (and K slice).  The K chunks of consecutive jobs form one
    // stream: the chunks of the next job are loaded while the current job's
    // last chunks are multiplied and its C tile written, so loads stay in
    // flight instead of draining at every job.
    //
    // Returns whether the stream stored an Inf or NaN that its caller is to
    // recompute (split modes, sources with REPAIR == kRepairByCaller).
    template <class Src>
    __device__ __forceinline__ bool stream(const Src &src, int first, int stride, char *lds)
    {
        const int tid = threadIdx.x;
        const int lane = tid & (kWave - 1);
        const int wave = uniform(tid >> 6);
        if (first >= src.count()) return false;
        const GemmDesc d0 = src.get(first);
        Cursor cur;
        cur.job = first;
        cur.c = 0;
        cur.nch = nchunks(d0.K);
        Cursor n1 = cur;
        advance(n1, src, stride);
        Cursor n2 = n1;
        advance(n2, src, stride);
        StageA a0, a1;
        StageBFor<Src> b0, b1;
        const auto px = lanes(src, d0, tid);
        load_chunk(a0, b0, src, px, d0, 0, true, tid);
        load_chunk(a1, b1, src, px, src.get(max(n1.job, 0)), n1.c * BK, n1.job >= 0, tid);
        begin(d0.M, d0.N, 0, 0, wave);
        if constexpr (Src::IM2COL == kIm2colTR) b0.fix_runs();
        chunk_to_lds(a0, b0, lds_buf(lds, 0), tid);
        __syncthreads();
        int bad_lo = 0x7fffffff, bad_hi = -1;  // jobs of the stream holding an Inf / NaN (split modes)
        for (;;) {
            if (!step(cur, n1, n2, a0, b0, a1, b1, px, lds, 0, src, stride, tid, lane, wave, bad_lo, bad_hi)) break;
            if (!step(cur, n1, n2, a1, b1, a0, b0, px, lds, 1, src, stride, tid, lane, wave, bad_lo, bad_hi)) break;
        }
        if constexpr (Cfg::SPLIT && Src::REPAIR == kRepairAfterCall) {
            if (__builtin_expect(bad_hi >= 0, 0)) rescan_call(src, bad_lo, bad_hi, stride);
        }
        return Cfg::SPLIT && Src::REPAIR == kRepairByCaller && bad_hi >= 0;
    }

    // IEEE special values in the split modes.  The bf16 pieces of an Inf
    // are not Infs (Inf - Inf makes the later pieces NaN), a finite value
    // above the largest bf16 rounds its first piece to Inf, and a cross
    // product of an Inf with a zero or opposite-signed later piece of the
    // other operand is NaN where FP32's Inf * b is Inf.  In the other
    // direction no special value is lost: an Inf or NaN operand meets every
    // piece of its partner, one of which is nonzero unless the partner is 0
    // or below bf16's range (NaN either way), so an element that FP32 makes
    // Inf or NaN comes out Inf or NaN here too, up to the summation order
    // near overflow, which FP32 implementations do not fix either.  So once
    // a job's sums are complete, the elements that are not finite are left
    // out of the tile's store and recomputed in FP32 afterwards (each by its
    // lane alone, FMA over k in order); the others are kept.  Detection is
    // one FMA per accumulator (x * 0 + s stays 0 unless x is Inf or NaN)
    // and a wave-wide test.

    // Whether any accumulator of the wave is Inf or NaN (uniform).
    __device__ __forceinline__ bool any_nonfinite() const
    {
        if (!live) return false;
        float s[2] = {0.f, 0.f};
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
            for (int b = 0; b < WN; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) s[i & 1] = __builtin_fmaf(acc[a][b][i], 0.f, s[i & 1]);
        const float t = s[0] + s[1];
        return __builtin_amdgcn_ballot_w64(t != t) != 0;
    }

    // The lane's Inf / NaN accumulators, bit (a * WN + b) * 4 + i (an
    // all-ones exponent, tested on the bits: a float compare per element
    // raised the kernels' register counts).
    __device__ __forceinline__ unsigned long long nonfinite_mask() const
    {
        static_assert(WM * WN * 4 <= 64, "one mask bit per accumulator");
        unsigned lo = 0, hi = 0;
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
            for (int b = 0; b < WN; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int e = (a * WN + b) * 4 + i;
                    // exponent + 1 reaches 0x100 only for an all-ones exponent
                    const unsigned x = (((__builtin_bit_cast(unsigned, acc[a][b][i]) >> 23) & 0xffu) + 1u) >> 8;
                    if (e < 32)
                        lo |= x << (e & 31);
                    else
                        hi |= x << (e & 31);
                }
        return ((unsigned long long)hi << 32) | lo;
    }

    // The FP32 sum over k of element (row, col) of job `job` and the jobs
    // chained into it, one FMA at a time in k order.
    template <class Src>
    __device__ __forceinline__ static float dot_fp32(const Src &src, int j0, int job, int stride, int row, int col)
    {
        float v = 0.f;
        for (int j = j0; j <= job; j += stride) {
            const GemmDesc d = src.get(j);
            const float *A = (const float *)d.A;
#pragma clang loop unroll(disable)
            for (int k = 0; k < d.K; ++k) {
                const float x = Cfg::TA ? A[(long long)k * d.lda + row] : A[(long long)row * d.lda + k];
                v = __builtin_fmaf(x, b_elem(src, d, k, col), v);
            }
        }
        return v;
    }

    // The first job of the chain that ends with `job`.
    template <class Src>
    __device__ __forceinline__ static int chain_start(const Src &src, int job, int stride)
    {
        int j0 = job;
        while (j0 - stride >= 0 && (src.get(j0 - stride).flags & kChain) != 0) j0 -= stride;
        return j0;
    }

    // Jobs lo, lo + stride, ..., hi of a stream, already stored: this wave's
    // elements of their C tiles that are Inf or NaN are recomputed in FP32.
    // The wave reads back what it stored itself.
    template <class Src>
    __device__ __forceinline__ static void rescan(const Src &src, int lo, int hi, int stride)
    {
        const auto cm = src.cols();
        const int lane = threadIdx.x & (kWave - 1);
        const int r = lane & 15, q = lane >> 4;
        // the wave's block (as layout() places it), derived afresh so that
        // nothing stays live across the stream for this rare path
        const int wave = uniform(threadIdx.x >> 6);
        const int wr0 = Cfg::wave_g(wave), wc0 = Cfg::wave_h(wave);
        for (int job = lo; job <= hi; job += stride) {
            const GemmDesc dj = src.get(job);
            if (dj.flags & kChain) continue;  // stored with the chain's last job
            const int j0 = chain_start(src, job, stride);
#pragma clang loop unroll(disable)
            for (int e = 0; e < WM * WN * 4; ++e) {
                const int ab = e >> 2, a = ab / WN, b = ab - a * WN;
                const int row = Cfg::row_block(wr0, a) * 16 + acc_row(q, e & 3);
                const int col = Cfg::col_block(wc0, b) * 16 + r;
                if (row >= dj.M || col >= dj.N) continue;
                float *o = (float *)dj.C + (long long)row * dj.ldc + cm(col);
                if (!__builtin_isfinite(*o)) *o = dot_fp32(src, j0, job, stride, row, col);
            }
        }
    }

    // rescan() out of line, for the sources whose stream loop it made
    // take more registers when inlined (nothing of the loop is live where
    // it runs, but the allocation of the whole kernel changed).
    template <class Src>
    __device__ __attribute__((noinline)) static void rescan_call(const Src src, int lo, int hi, int stride)
    {
        rescan(src, lo, hi, stride);
    }

    // The elements of mask `bad` (see nonfinite_mask()) of the job just stored
    // without them, recomputed in FP32 and written as store() writes them.
    // The FMA chain runs over the job and the jobs chained into it.
    template <class Src, class Cols>
    __device__ __forceinline__ void repair(const Src &src, int job, int stride, unsigned long long bad,
                                           const Cols &cm) const
    {
        const int j0 = chain_start(src, job, stride);
        const GemmDesc dj = src.get(job);
        const int lane = threadIdx.x & (kWave - 1);
        const int r = lane & 15, q = lane >> 4;
        while (bad != 0) {
            const int e = __builtin_ctzll(bad);
            bad &= bad - 1;
            const int ab = e >> 2, a = ab / WN, b = ab - a * WN;
            const int row = Cfg::row_block(wr0, a) * 16 + acc_row(q, e & 3);
            const int col = Cfg::col_block(wc0, b) * 16 + r;
            if (row >= dj.M || col >= dj.N) continue;
            float v = dot_fp32(src, j0, job, stride, row, col);
            float *o = (float *)dj.C + (long long)row * dj.ldc + cm(col);
            if (scale && !Cols::MASK_ROWS) v = beta != T(0) ? alpha * v + beta * *o : alpha * v;
            *o = v;
        }
    }

    // Element (k, col) of job d's right operand, k and col relative to the
    // job: stored (plain or transposed), or gathered from a convolution's
    // input as the job's loads do (see Stage::load_im2col / load_im2col_t).
    template <class Src>
    __device__ __forceinline__ static float b_elem(const Src &src, const GemmDesc &d, int k, int col)
    {
        const float *B = (const float *)d.B;
        if constexpr (Src::IM2COL == kPlainB) {
            return Cfg::TB ? B[(long long)col * d.ldb + k] : B[(long long)k * d.ldb + col];
        } else {
            const ConvGeom &g = src.geom;
            int t, n, off = 0;
            if constexpr (Src::IM2COL == kIm2col) {
                t = k;             // tap from the job's first channel
                n = d.ldb + col;   // output pixel (the job's ldb is the tile's first)
                if (g.fold) {
                    const int img = udiv(n, g.mPQ);
                    n -= img * g.PQ;
                    off = img * g.img;
                }
            } else {
                t = src.c0 + col;  // the stored row is the tap
                n = src.rho0 + k;  // k runs over output pixels
            }
            // the divisions by the geometry's multipliers (a division by a
            // variable would put its reciprocal setup in registers)
            const int c = udiv(t, g.mRS), rem = t - c * g.RS, fr = udiv(rem, g.mS), fs = rem - fr * g.S;
            const int p = udiv(n, g.mQ), pq = n - p * g.Q;
            int ih = p * g.sh - g.ph + fr * g.dh, iw = pq * g.sw - g.pw + fs * g.dw;
            if (g.th != 1 || g.tw != 1) {
                // floor division (ih, iw >= -2^13 > -1024 * stride)
                const int a = udiv(ih + 1024 * g.th, g.mTh) - 1024, b = udiv(iw + 1024 * g.tw, g.mTw) - 1024;
                if (ih != a * g.th || iw != b * g.tw) return 0.f;
                ih = a;
                iw = b;
            }
            if ((unsigned)ih >= (unsigned)g.H || (unsigned)iw >= (unsigned)g.W) return 0.f;
            return B[off + ((long long)c * g.H + ih) * g.W + iw];
        }
    }

    // Row of the 16x16 sub-tile held in accumulator register i of lane group
    // q.  v_mfma_f32_16x16x4_f32 gives lane (r, q) rows 4q..4q+3;
    // v_mfma_f64_16x16x4_f64 interleaves them: rows q, q+4, q+8, q+12
    // (column r in both; measured on gfx950).
    __device__ __forceinline__ static int acc_row(int q, int i)
    {
        if constexpr (sizeof(T) == 8)
            return q + 4 * i;
        else
            return 4 * q + i;
    }

    // Store the accumulators to a row-major OutT matrix (a C tile or a split-K
    // partial) with leading dimension ldo.  The buffer ends right after
    // element (M-1, N-1), so rows >= M fall past it; columns >= N (in range,
    // they belong to the next tile) take an out-of-range offset, and so do
    // the elements of mask `skip` (bit (a * WN + b) * 4 + i).
    template <typename OutT, class Cols = PlainCols>
    __device__ __forceinline__ void store(OutT *__restrict__ out, int ldo, int M, int N,
                                          const Cols &cm = Cols{}, unsigned long long skip = 0) const
    {
        constexpr int ES_OUT = sizeof(OutT);
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(out, cm.extent(M, ldo, N) * ES_OUT);
        const int lane = threadIdx.x & (kWave - 1);
        const int r = lane & 15, q = lane >> 4;
        const int rstep = (acc_row(0, 1) - acc_row(0, 0)) * ldo * ES_OUT;  // offset between a lane's rows
        if (!live) return;
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
            for (int b = 0; b < WN; ++b) {
                // FP32 / FP64 skip the sub-tiles past the job's rows or
                // columns (a prefix of the wave's; uniform branches)
                if (Cfg::INTERLEAVE && (a >= wm || b >= wn)) continue;
                const int col = Cfg::col_block(wc0, b) * 16 + r;
                const int row = Cfg::row_block(wr0, a) * 16 + acc_row(q, 0);
                const int off = (row * ldo + cm(col)) * ES_OUT | (col < N ? 0 : kOOB);
                // element i's offset, out of range when `skip` leaves it out
                auto at = [&](int i) { return (skip >> ((a * WN + b) * 4 + i)) & 1 ? kOOB : off + i * rstep; };
                if constexpr (Cols::MASK_ROWS) {
                    // rows past M masked one by one (no alpha / beta here)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int ri = Cfg::row_block(wr0, a) * 16 + acc_row(q, i);
                        buf_store1<OutT>(rs, ri < M ? at(i) : kOOB, (OutT)acc[a][b][i]);
                    }
                } else if (scale) {
                    OutT old[4] = {OutT(0), OutT(0), OutT(0), OutT(0)};
                    if (beta != T(0)) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) old[i] = buf_load1<OutT>(rs, off + i * rstep);
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        buf_store1<OutT>(rs, at(i),
                                         beta != T(0) ? (OutT)(alpha * acc[a][b][i] + beta * old[i])
                                                      : (OutT)(alpha * acc[a][b][i]));
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i) buf_store1<OutT>(rs, at(i), (OutT)acc[a][b][i]);
                }
            }
    }
};

// Job source over a device array of jobs.
struct ArraySrc {
    static constexpr int IM2COL = kPlainB;
    static constexpr int REPAIR = kRepairAfterCall;
    const GemmDesc *__restrict__ jobs;
    int n;
    __device__ __forceinline__ PlainCols cols() const { return {}; }
    __device__ __forceinline__ int count() const { return n; }
    // The descriptor is read through the constant address space with a
    // uniform index, so it arrives by scalar loads (lgkmcnt), not vector
    // loads that would count against the operand prefetches.
    __device__ __forceinline__ GemmDesc get(int j) const
    {
#if __HIP_DEVICE_COMPILE__
        typedef const __attribute__((address_space(4))) GemmDesc *cdesc;
        return ((cdesc)jobs)[uniform(j)];
#else
        return jobs[j];  // host pass only parses device functions
#endif
    }
};

// Job source of exactly one job held in registers.
struct OneSrc {
    static constexpr int IM2COL = kPlainB;
    static constexpr int REPAIR = kRepairAtStore;
    GemmDesc job;
    __device__ __forceinline__ PlainCols cols() const { return {}; }
    __device__ __forceinline__ int count() const { return 1; }
    __device__ __forceinline__ GemmDesc get(int) const { return job; }
};

// The tile (m0, n0) and K slice [kbeg, kend) of g as a job of its own, its
// C tile written to out (leading dimension ldo) at (m0, n0).
template <class Cfg>
__device__ __forceinline__ GemmDesc tile_job(const GemmDesc &g, int m0, int n0, int kbeg, int kend, void *out,
                                             int ldo)
{
    typedef typename Cfg::T T;
    GemmDesc j;
    const T *A = (const T *)g.A;
    const T *B = (const T *)g.B;
    j.A = A + (Cfg::TA ? (size_t)kbeg * g.lda + m0 : (size_t)m0 * g.lda + kbeg);
    j.B = B + (Cfg::TB ? (size_t)n0 * g.ldb + kbeg : (size_t)kbeg * g.ldb + n0);
    j.C = (T *)out + (size_t)m0 * ldo + n0;
    j.M = min(Cfg::BM, g.M - m0);
    j.N = min(Cfg::BN, g.N - n0);
    j.K = max(0, kend - kbeg);
    j.lda = g.lda;
    j.ldb = g.ldb;
    j.ldc = ldo;
    j.flags = g.flags;
    return j;
}

}  // namespace ls
