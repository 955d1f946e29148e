// gemm_core.hpp -- the tile GEMM at the heart of every LibShalom entry point,
// written for gfx950 (CDNA4): 64-wide wavefronts, MFMA matrix cores, LDS.
//
// One workgroup (4 waves, 256 threads) computes one BM x BN tile of
// C = op(A) * op(B) (row-major, beta = 0 overwrite) over a K range.
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
// * Register tiling.  Each wave owns a rectangle of 16x16 sub-tiles (at most
//   BM/32 x BN/32); the wave grid (2x2, 1x4 or 4x1) is chosen per tile at run
//   time so small or skinny tiles keep all four waves busy.
// * Pipelining.  The next K chunk is loaded into registers while the current
//   one is multiplied out of LDS; LDS is double-buffered when it fits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ls {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

enum Prec : int { P_FP32 = 0, P_BF16X3 = 1, P_BF16 = 2, P_BF16X6 = 3, P_FP64 = 4 };

constexpr int kThreads = 256;
constexpr int kWave = 64;

// One GEMM.  Leading dimensions are in elements of the storage.
struct GemmDesc {
    const void *A;
    const void *B;
    void *C;
    int M, N, K;
    int lda, ldb, ldc;
    int flags;  // bit 0: A 16-byte vector loads allowed, bit 1: B
};

// Work item of a batched launch: which GEMM and which tile of it.
struct WorkItem {
    int gemm;
    int tile;  // tm * tiles_n + tn
};

template <int PREC> struct PrecInfo;
template <> struct PrecInfo<P_FP32> { typedef float T; static constexpr int planes = 0; };
template <> struct PrecInfo<P_FP64> { typedef double T; static constexpr int planes = 0; };
template <> struct PrecInfo<P_BF16> { typedef float T; static constexpr int planes = 1; };
template <> struct PrecInfo<P_BF16X3> { typedef float T; static constexpr int planes = 2; };
template <> struct PrecInfo<P_BF16X6> { typedef float T; static constexpr int planes = 3; };

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Split x into bf16 pieces, x ~= p0 + p1 + p2, each round-to-nearest-even of
// the remaining residual (the residual subtraction is exact in FP32).
template <int NP>
__device__ __forceinline__ void split_bf16(float x, __bf16 (&p)[3])
{
    __bf16 h = (__bf16)x;
    p[0] = h;
    if constexpr (NP >= 2) {
        float r = x - (float)h;
        __bf16 l = (__bf16)r;
        p[1] = l;
        if constexpr (NP >= 3) {
            float r2 = r - (float)l;
            p[2] = (__bf16)r2;
        }
    }
}

// ---------------------------------------------------------------------------
// Configuration of one tile kernel.
// ---------------------------------------------------------------------------
template <int PREC_, int TA_, int TB_, int BM_, int BN_>
struct TileCfg {
    static constexpr int PREC = PREC_, TA = TA_, TB = TB_, BM = BM_, BN = BN_;
    typedef typename PrecInfo<PREC>::T T;
    static constexpr int NP = PrecInfo<PREC>::planes;  // 0: no split (FP32/FP64 MFMA)
    static constexpr bool SPLIT = NP > 0;
    static constexpr int BK = sizeof(T) == 8 ? 16 : 32;
    static constexpr int VEC = 16 / sizeof(T);            // elements per 16-byte load
    static constexpr int NSLOT = BK / 4;                  // 16x16x4 MFMA k-steps per chunk
    // LDS panels.  Unsplit: [rows][LDK] of T, k position permuted (see kpos).
    // Split: [plane][rows][LDKH] of bf16, natural k order.
    static constexpr int LDK = BK + VEC;                  // 16-byte row pad
    static constexpr int LDKH = BK + 8;                   // 16-byte row pad (bf16)
    static constexpr int A_BYTES = SPLIT ? NP * BM * LDKH * 2 : BM * LDK * (int)sizeof(T);
    static constexpr int B_BYTES = SPLIT ? NP * BN * LDKH * 2 : BN * LDK * (int)sizeof(T);
    static constexpr int BUF_BYTES = A_BYTES + B_BYTES;
    static constexpr int NBUF = (2 * BUF_BYTES <= 80 * 1024) ? 2 : 1;
    static constexpr int LDS_BYTES = NBUF * BUF_BYTES;
    // Workgroups per CU the LDS allows (160 KB per CU); the persistent kernel
    // asks the compiler to fit its registers to this residency.
    static constexpr int WG_PER_CU = (160 * 1024 / LDS_BYTES) < 8 ? (160 * 1024 / LDS_BYTES) : 8;
    static constexpr int WM_MAX = BM / 32, WN_MAX = BN / 32;  // sub-tiles per wave (2x2 grid fits)
    static constexpr bool A_KCONTIG = !TA;  // A[m][k] rows contiguous along k
    static constexpr bool B_KCONTIG = TB;   // B^T[n][k] contiguous along k
    typedef typename std::conditional<sizeof(T) == 8, f64x4, f32x4>::type Acc;
    static_assert(BM % 32 == 0 && BN % 32 == 0, "tile must be a multiple of 32");
    static_assert(WM_MAX * WN_MAX <= 16, "at most 16 sub-tiles per wave");
};

// LDS position of k inside a chunk for the unsplit layout: lane group q of a
// 16x16x4 MFMA takes k = 4*s + q at step s, and a lane reads its NSLOT values
// for steps 0..NSLOT-1 as one contiguous run, so k = 4*s + q is stored at
// q*NSLOT + s.  Steps whose k are all >= K can then be skipped.
template <int NSLOT>
__device__ __forceinline__ int kpos(int k) { return (k & 3) * NSLOT + (k >> 2); }

// ---------------------------------------------------------------------------
// Global -> register staging of one operand panel chunk (ROWS x BK).
// "rows" are M for A and N for B.  KCONTIG: storage is [row][k] (stride ld);
// otherwise [k][row].
// ---------------------------------------------------------------------------
template <typename T, int ROWS, int BK, bool KCONTIG>
struct Stage {
    static constexpr int VEC = 16 / sizeof(T);
    static constexpr int NVEC = ROWS * BK / VEC;
    static constexpr int PER = (NVEC + kThreads - 1) / kThreads;
    T v[PER][VEC];

    __device__ __forceinline__ static void coords(int idx, int &row, int &k)
    {
        if constexpr (KCONTIG) {
            row = idx / (BK / VEC);
            k = (idx % (BK / VEC)) * VEC;
        } else {
            k = idx / (ROWS / VEC);
            row = (idx % (ROWS / VEC)) * VEC;
        }
    }

    __device__ __forceinline__ void load(const T *__restrict__ base, int ld, int R, int kend,
                                         int row0, int k0, bool vec_ok, int tid)
    {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = tid + i * kThreads;
            int row, k;
            coords(idx, row, k);
            const int gr = row0 + row, gk = k0 + k;
#pragma unroll
            for (int e = 0; e < VEC; ++e) v[i][e] = T(0);
            const bool active = (NVEC % kThreads == 0) || idx < NVEC;
            if constexpr (KCONTIG) {
                if (active && gr < R) {
                    const T *p = base + (size_t)gr * ld + gk;
                    if (vec_ok && gk + VEC <= kend) {
                        if constexpr (sizeof(T) == 4) {
                            f32x4 t = *(const f32x4 *)p;
#pragma unroll
                            for (int e = 0; e < VEC; ++e) v[i][e] = t[e];
                        } else {
                            f64x2 t = *(const f64x2 *)p;
#pragma unroll
                            for (int e = 0; e < VEC; ++e) v[i][e] = t[e];
                        }
                    } else {
#pragma unroll
                        for (int e = 0; e < VEC; ++e)
                            if (gk + e < kend) v[i][e] = p[e];
                    }
                }
            } else {
                if (active && gk < kend) {
                    const T *p = base + (size_t)gk * ld + gr;
                    if (vec_ok && gr + VEC <= R) {
                        if constexpr (sizeof(T) == 4) {
                            f32x4 t = *(const f32x4 *)p;
#pragma unroll
                            for (int e = 0; e < VEC; ++e) v[i][e] = t[e];
                        } else {
                            f64x2 t = *(const f64x2 *)p;
#pragma unroll
                            for (int e = 0; e < VEC; ++e) v[i][e] = t[e];
                        }
                    } else {
#pragma unroll
                        for (int e = 0; e < VEC; ++e)
                            if (gr + e < R) v[i][e] = p[e];
                    }
                }
            }
        }
    }

    // Unsplit LDS layout: [ROWS][LDK] with permuted k.
    template <int LDK>
    __device__ __forceinline__ void store_plain(T *__restrict__ lds, int tid) const
    {
        constexpr int NSLOT = BK / 4;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = tid + i * kThreads;
            if (!((NVEC % kThreads == 0) || idx < NVEC)) continue;
            int row, k;
            coords(idx, row, k);
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                if constexpr (KCONTIG)
                    lds[row * LDK + kpos<NSLOT>(k + e)] = v[i][e];
                else
                    lds[(row + e) * LDK + kpos<NSLOT>(k)] = v[i][e];
            }
        }
    }

    // Split LDS layout: NP planes of [ROWS][LDKH] bf16, natural k order.
    template <int NP, int LDKH>
    __device__ __forceinline__ void store_split(__bf16 *__restrict__ lds, int tid) const
    {
        static_assert(sizeof(T) == 4, "split modes take FP32 inputs");
        constexpr int PLANE = ROWS * LDKH;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = tid + i * kThreads;
            if (!((NVEC % kThreads == 0) || idx < NVEC)) continue;
            int row, k;
            coords(idx, row, k);
            __bf16 p[VEC][3];
#pragma unroll
            for (int e = 0; e < VEC; ++e) split_bf16<NP>(v[i][e], p[e]);
            if constexpr (KCONTIG) {
                // four consecutive k of one row: one 8-byte write per plane
#pragma unroll
                for (int pl = 0; pl < NP; ++pl) {
                    bf16x4 w = {p[0][pl], p[1][pl], p[2][pl], p[3][pl]};
                    *(bf16x4 *)(lds + pl * PLANE + row * LDKH + k) = w;
                }
            } else {
#pragma unroll
                for (int pl = 0; pl < NP; ++pl)
#pragma unroll
                    for (int e = 0; e < VEC; ++e) lds[pl * PLANE + (row + e) * LDKH + k] = p[e][pl];
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
    static constexpr int WM = Cfg::WM_MAX, WN = Cfg::WN_MAX;
    typedef Stage<T, BM, BK, Cfg::A_KCONTIG> StageA;
    typedef Stage<T, BN, BK, Cfg::B_KCONTIG> StageB;

    Acc acc[WM][WN];
    int wr0, wc0, wm, wn;  // this wave's sub-tile rectangle (in 16-row/col units)

    // Pick the wave grid for a tile whose valid extent is mt x nt sub-tiles.
    __device__ __forceinline__ void layout(int mt, int nt, int wave)
    {
        // candidates: 2x2, 1x4 (needs mt <= WM), 4x1 (needs nt <= WN)
        int lm = 2, ln = 2;
        int pm = (mt + 1) >> 1, pn = (nt + 1) >> 1;
        int best = pm * pn;
        if (mt <= WM) {
            int q = (nt + 3) >> 2;
            if (mt * q < best) { best = mt * q; lm = 1; ln = 4; pm = mt; pn = q; }
        }
        if (nt <= WN) {
            int q = (mt + 3) >> 2;
            if (q * nt < best) { best = q * nt; lm = 4; ln = 1; pm = q; pn = nt; }
        }
        const int wi = wave / ln, wj = wave % ln;
        wr0 = wi * pm;
        wc0 = wj * pn;
        wm = min(pm, mt - wr0);
        wn = min(pn, nt - wc0);
        if (wm < 0) wm = 0;
        if (wn < 0) wn = 0;
        (void)lm;
    }

    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
            for (int b = 0; b < WN; ++b) acc[a][b] = Acc{0, 0, 0, 0};
    }

    // Multiply one LDS chunk.  ksteps = number of valid 4-wide k steps (unsplit).
    __device__ __forceinline__ void compute(const char *__restrict__ buf, int lane, int ksteps)
    {
        const int r = lane & 15, q = lane >> 4;
        if constexpr (!Cfg::SPLIT) {
            constexpr int NS = Cfg::NSLOT, LDK = Cfg::LDK;
            const T *As = (const T *)buf;
            const T *Bs = (const T *)(buf + Cfg::A_BYTES);
            T af[WM][NS], bfr[WN][NS];
#pragma unroll
            for (int a = 0; a < WM; ++a) {
                if (a < wm) {
                    const T *p = As + ((wr0 + a) * 16 + r) * LDK + q * NS;
#pragma unroll
                    for (int s = 0; s < NS; s += 16 / sizeof(T)) {
                        if constexpr (sizeof(T) == 4) {
                            f32x4 t = *(const f32x4 *)(p + s);
                            af[a][s] = t[0]; af[a][s + 1] = t[1]; af[a][s + 2] = t[2]; af[a][s + 3] = t[3];
                        } else {
                            f64x2 t = *(const f64x2 *)(p + s);
                            af[a][s] = t[0]; af[a][s + 1] = t[1];
                        }
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < WN; ++b) {
                if (b < wn) {
                    const T *p = Bs + ((wc0 + b) * 16 + r) * LDK + q * NS;
#pragma unroll
                    for (int s = 0; s < NS; s += 16 / sizeof(T)) {
                        if constexpr (sizeof(T) == 4) {
                            f32x4 t = *(const f32x4 *)(p + s);
                            bfr[b][s] = t[0]; bfr[b][s + 1] = t[1]; bfr[b][s + 2] = t[2]; bfr[b][s + 3] = t[3];
                        } else {
                            f64x2 t = *(const f64x2 *)(p + s);
                            bfr[b][s] = t[0]; bfr[b][s + 1] = t[1];
                        }
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (s < ksteps) {
#pragma unroll
                    for (int a = 0; a < WM; ++a)
#pragma unroll
                        for (int b = 0; b < WN; ++b)
                            if (a < wm && b < wn) {
                                if constexpr (sizeof(T) == 4)
                                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a][s], bfr[b][s], acc[a][b], 0, 0, 0);
                                else
                                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a][s], bfr[b][s], acc[a][b], 0, 0, 0);
                            }
                }
            }
        } else {
            constexpr int NP = Cfg::NP, LDKH = Cfg::LDKH;
            const __bf16 *As = (const __bf16 *)buf;
            const __bf16 *Bs = (const __bf16 *)(buf + Cfg::A_BYTES);
            bf16x8 af[WM][NP], bfr[WN][NP];
#pragma unroll
            for (int a = 0; a < WM; ++a)
                if (a < wm)
#pragma unroll
                    for (int pl = 0; pl < NP; ++pl)
                        af[a][pl] = *(const bf16x8 *)(As + pl * BM * LDKH + ((wr0 + a) * 16 + r) * LDKH + q * 8);
#pragma unroll
            for (int b = 0; b < WN; ++b)
                if (b < wn)
#pragma unroll
                    for (int pl = 0; pl < NP; ++pl)
                        bfr[b][pl] = *(const bf16x8 *)(Bs + pl * BN * LDKH + ((wc0 + b) * 16 + r) * LDKH + q * 8);
#pragma unroll
            for (int a = 0; a < WM; ++a)
#pragma unroll
                for (int b = 0; b < WN; ++b)
                    if (a < wm && b < wn) {
                        Acc c = acc[a][b];
                        // smallest partial products first
                        if constexpr (NP >= 3) {
                            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][2], bfr[b][0], c, 0, 0, 0);
                            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][1], bfr[b][1], c, 0, 0, 0);
                            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][0], bfr[b][2], c, 0, 0, 0);
                        }
                        if constexpr (NP >= 2) {
                            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][1], bfr[b][0], c, 0, 0, 0);
                            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][0], bfr[b][1], c, 0, 0, 0);
                        }
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a][0], bfr[b][0], c, 0, 0, 0);
                        acc[a][b] = c;
                    }
        }
    }

    template <class StageT>
    __device__ __forceinline__ static void store_stage(const StageT &st, char *dst, int tid)
    {
        if constexpr (!Cfg::SPLIT)
            st.template store_plain<Cfg::LDK>((T *)dst, tid);
        else
            st.template store_split<Cfg::NP, Cfg::LDKH>((__bf16 *)dst, tid);
    }

    // Wave layout for a tile of d at (m0, n0), accumulators cleared.
    __device__ __forceinline__ void begin(int M, int N, int m0, int n0, int wave)
    {
        layout((min(BM, M - m0) + 15) >> 4, (min(BN, N - n0) + 15) >> 4, wave);
        zero();
    }

    // Global -> registers: the K chunk at k0 of the tile at (m0, n0).
    __device__ __forceinline__ static void load_chunk(StageA &sa, StageB &sb, const GemmDesc &d, int m0, int n0,
                                                      int k0, int kend, int tid)
    {
        sa.load((const T *)d.A, d.lda, d.M, kend, m0, k0, d.flags & 1, tid);
        sb.load((const T *)d.B, d.ldb, d.N, kend, n0, k0, d.flags & 2, tid);
    }

    // Registers -> LDS in the packed MFMA operand layout (transpose / split).
    __device__ __forceinline__ static void chunk_to_lds(const StageA &sa, const StageB &sb, char *buf, int tid)
    {
        store_stage(sa, buf, tid);
        store_stage(sb, buf + Cfg::A_BYTES, tid);
    }

    __device__ __forceinline__ static int ksteps(int kend, int k0) { return min(Cfg::NSLOT, (kend - k0 + 3) >> 2); }

    __device__ __forceinline__ static char *lds_buf(char *lds, int i)
    {
        return lds + (Cfg::NBUF == 2 ? (i & 1) * Cfg::BUF_BYTES : 0);
    }

    // Run the K loop over [kbeg, kend) for the tile at (m0, n0) of d.
    __device__ __forceinline__ void run(const GemmDesc &d, int m0, int n0, int kbeg, int kend, char *lds)
    {
        const int tid = threadIdx.x;
        const int lane = tid & (kWave - 1);
        const int wave = uniform(tid >> 6);
        begin(d.M, d.N, m0, n0, wave);
        if (kend <= kbeg) return;
        const int nchunks = (kend - kbeg + BK - 1) / BK;
        StageA sa;
        StageB sb;
        load_chunk(sa, sb, d, m0, n0, kbeg, kend, tid);
        chunk_to_lds(sa, sb, lds, tid);
        __syncthreads();
        for (int c = 0; c < nchunks; ++c) {
            const int k0 = kbeg + c * BK;
            const bool more = c + 1 < nchunks;
            if (more) load_chunk(sa, sb, d, m0, n0, k0 + BK, kend, tid);
            compute(lds_buf(lds, c), lane, ksteps(kend, k0));
            if (more) {
                if constexpr (Cfg::NBUF == 1) __syncthreads();
                chunk_to_lds(sa, sb, lds_buf(lds, c + 1), tid);
            }
            __syncthreads();
        }
    }

    // Persistent job stream: this workgroup takes jobs[first], jobs[first +
    // stride], ...  Each job is a GEMM of at most BM x BN whose pointers are
    // already offset to the tile.  The stream of K chunks runs across job
    // boundaries: the first chunk of the next job is loaded while the last
    // chunk of the current one is multiplied and its C tile written, so the
    // loads stay in flight instead of draining at every job.
    __device__ __forceinline__ void stream_jobs(const GemmDesc *__restrict__ jobs, int njobs, int first, int stride,
                                                char *lds)
    {
        const int tid = threadIdx.x;
        const int lane = tid & (kWave - 1);
        const int wave = uniform(tid >> 6);
        int j = first;
        if (j >= njobs) return;
        GemmDesc d = jobs[j];
        int nj = j + stride;
        GemmDesc nd = d;
        if (nj < njobs) nd = jobs[nj];
        begin(d.M, d.N, 0, 0, wave);
        int nch = max(1, (d.K + BK - 1) / BK), c = 0, buf = 0;
        StageA sa;
        StageB sb;
        load_chunk(sa, sb, d, 0, 0, 0, d.K, tid);
        chunk_to_lds(sa, sb, lds, tid);
        __syncthreads();
        for (;;) {
            const bool last = c + 1 == nch;
            const bool more = !last || nj < njobs;
            if (more) {
                if (!last)
                    load_chunk(sa, sb, d, 0, 0, (c + 1) * BK, d.K, tid);
                else
                    load_chunk(sa, sb, nd, 0, 0, 0, nd.K, tid);
            }
            compute(lds_buf(lds, buf), lane, ksteps(d.K, c * BK));
            if (last) store((T *)d.C, d.ldc, d.M, d.N, 0, 0);
            if (!more) break;
            if (last) {
                d = nd;
                nj += stride;
                if (nj < njobs) nd = jobs[nj];
                begin(d.M, d.N, 0, 0, wave);
                nch = max(1, (d.K + BK - 1) / BK);
                c = 0;
            } else {
                ++c;
            }
            if constexpr (Cfg::NBUF == 1) __syncthreads();
            chunk_to_lds(sa, sb, lds_buf(lds, buf + 1), tid);
            __syncthreads();
            ++buf;
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

    // Store the accumulators to a row-major OutT matrix (C or a split-K
    // partial) with leading dimension ldo; rows >= M / cols >= N are masked.
    template <typename OutT>
    __device__ __forceinline__ void store(OutT *__restrict__ out, int ldo, int M, int N, int m0, int n0) const
    {
        const int lane = threadIdx.x & (kWave - 1);
        const int r = lane & 15, q = lane >> 4;
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
            for (int b = 0; b < WN; ++b)
                if (a < wm && b < wn) {
                    const int col = n0 + (wc0 + b) * 16 + r;
                    const int row0 = m0 + (wr0 + a) * 16;
                    if (col < N) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int row = row0 + acc_row(q, i);
                            if (row < M) out[(size_t)row * ldo + col] = (OutT)acc[a][b][i];
                        }
                    }
                }
    }
};

// XCD-aware block remap: the dispatcher places workgroup b on XCD b % 8, so
// give each XCD a contiguous run of logical blocks (neighbouring tiles share
// operand panels through that XCD's L2).
__device__ __forceinline__ int xcd_remap(int b, int nblocks)
{
    constexpr int X = 8;
    const int per = nblocks / X;
    const int xcd = b % X, local = b / X;
    if (local >= per) return b;  // the last nblocks % 8 blocks keep their place
    return xcd * per + local;
}

}  // namespace ls
