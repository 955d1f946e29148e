// skinny.hpp -- one GEMM with a short side: a tall C (N <= 64, M large) or a
// wide one (M <= 64, N large), the irregular shapes of sgemm_mp
// (SMM_thread.c:75-114 splits their long side over threads); FP32 MFMA or
// the split-BF16 modes.
//
// The tile kernel splits such a GEMM's K over workgroups, which then meet
// in a fixup (partial tiles written, an arrival counter, the last slice
// summing them: two more memory round trips), and each workgroup's K slice
// still arrives two chunks at a time.  Here a workgroup owns 16 indices of
// the long side over all of K: its four waves split K between them, each
// keeping D groups of 16 k in flight with loads straight into the MFMA
// operand registers (no LDS staging, no barrier in the K loop), and the
// four partial sums meet once, in LDS, before the store.  One launch, no
// scratch memory.
//
// Roles.  The streamed operand X supplies the MFMA's first operand (16
// rows: the workgroup's long-side indices), the small operand Y its second
// (16 columns per block, NC blocks covering the short side):
//   tall: X = op(A) (rows m), Y = op(B) (columns n); D[i][j] = C[m0+i][j]
//   wide: X = op(B)^T (rows n), Y = op(A)^T (columns m); D[i][j] = C[j][n0+i]
// Lane (r, q) of v_mfma_f32_16x16x4_f32 supplies X[r][k] and Y[k][r] of its
// k step and holds D rows 4q..4q+3 of column r.  Step s of a group at k0
// takes k = k0 + 4q + s, so a lane's four k of a group are consecutive: an
// operand stored k-contiguous (KC: element (index, k) at index*ld + k) is one
// 16-byte load per lane and group, an index-contiguous one (IC: at k*ld +
// index) four 4-byte loads, each covering 64 contiguous bytes per 16 lanes.
// The split modes take groups of 32 k, lane q's eight consecutive k (the
// v_mfma_f32_16x16x32_bf16 operand layout), split into bf16 pieces in
// registers.
#pragma once
#include "gemm_core.hpp"

namespace ls {

struct SkinnyArgs {
    const float *A, *B;
    float *C;
    int M, N, K, lda, ldb, ldc;
    float alpha, beta;
};

// Groups of k in flight per wave: more for the narrower small sides, whose
// groups are fewer bytes; half as many groups of 32 k in the split modes.
template <int NC, bool SPLIT>
constexpr int skinny_depth() { return (NC == 1 ? 8 : NC == 2 ? 6 : 4) / (SPLIT ? 2 : 1); }

// One operand's loads for a group: a lane's V consecutive k of its index
// (V = 4 for FP32's groups of 16 k, 8 for the split modes' groups of 32):
// KC, V / 4 16-byte loads per lane; IC, V 4-byte loads.  Elements past the
// wave's K range or past the operand's index count read 0.
template <bool KC, int V>
struct SkinnyOperand {
    __amdgpu_buffer_rsrc_t rs;
    int base[KC ? 1 : V];  // byte offset of the lane's k0 = 0 elements
    int kstep;             // bytes per k
    bool idx_ok;           // the lane's index is inside the operand

    // index i of this lane, k offset kq of the lane within a group
    __device__ __forceinline__ void init(const float *p, int count, int ld, int K, int i, int kq)
    {
        // exactly the operand's extent: an index past count, or k past K in
        // an index-contiguous operand, falls outside it
        const int nrec = (count > 0 && K > 0) ? (KC ? ((count - 1) * ld + K) : ((K - 1) * ld + count)) * 4 : 0;
        rs = make_rsrc(p, nrec);
        idx_ok = i < count;
        if constexpr (KC) {
            base[0] = (i * ld + kq) * 4;
            kstep = 4;
        } else {
#pragma unroll
            for (int s = 0; s < V; ++s) base[s] = ((kq + s) * ld + i) * 4;
            kstep = ld * 4;
        }
    }

    // Elements k0 + kq + s (s < V) of the lane's index; k >= kend reads 0
    // (the next wave's range, or past K: in range for a KC row).
    __device__ __forceinline__ void load(float (&v)[V], int k0, int kq, int kend) const
    {
        const int koff = k0 * kstep;
        if constexpr (KC) {
#pragma unroll
            for (int t = 0; t < V / 4; ++t) {
                float u[4];
                buf_load16<float>(rs, oob_unless(k0 < kend, base[0] + koff + 16 * t), u);
#pragma unroll
                for (int e = 0; e < 4; ++e) v[4 * t + e] = (k0 + kq + 4 * t + e < kend) ? u[e] : 0.f;
            }
        } else {
#pragma unroll
            for (int s = 0; s < V; ++s)
                v[s] = buf_load1<float>(rs, oob_unless(idx_ok & (k0 + kq + s < kend), base[s] + koff));
        }
    }

    // Element (i, k) read directly (the split modes' FP32 recompute).
    __device__ __forceinline__ static float at(const float *p, int ld, int i, int k)
    {
        return KC ? p[(long long)i * ld + k] : p[(long long)k * ld + i];
    }
};

// The split modes' pieces of a lane's eight consecutive k: plane pl as the
// bf16x8 operand of v_mfma_f32_16x16x32_bf16 (element e = k offset e).
template <int NP>
__device__ __forceinline__ void skinny_split(const float (&v)[8], bf16x8 (&pl)[NP])
{
    u32x2 lo[3], hi[3];
    const float a[4] = {v[0], v[1], v[2], v[3]}, b[4] = {v[4], v[5], v[6], v[7]};
    split4_bf16<NP>(a, lo);
    split4_bf16<NP>(b, hi);
#pragma unroll
    for (int p = 0; p < NP; ++p) pl[p] = __builtin_bit_cast(bf16x8, u32x4{lo[p][0], lo[p][1], hi[p][0], hi[p][1]});
}

template <int PREC, int TA, int TB, int WIDE, int NC, int NW>
__global__ __launch_bounds__(64 * NW) void k_gemm_skinny(SkinnyArgs a)
{
    constexpr int NP = PrecInfo<PREC>::planes;  // 0: FP32 MFMA
    constexpr bool SPLIT = NP > 0;
    constexpr int GK = SPLIT ? 32 : 16;  // k per group
    constexpr int V = GK / 4;            // consecutive k per lane
    constexpr int D = skinny_depth<NC, SPLIT>();
    // X: k-contiguous when op(A) is A (tall) or op(B)^T is B stored N x K (wide)
    constexpr bool XKC = WIDE ? TB != 0 : TA == 0;
    constexpr bool YKC = WIDE ? TA == 0 : TB != 0;
    __shared__ __attribute__((aligned(16))) float red[NW][16][16 * NC];

    const int lane = threadIdx.x & 63, wave = uniform((int)(threadIdx.x >> 6));
    const int r = lane & 15, q = lane >> 4, kq = V * q;
    const int L = WIDE ? a.N : a.M;  // long side
    const int S = WIDE ? a.M : a.N;  // short side
    // an index-contiguous X is read in 64-byte runs: the workgroups of two
    // neighbouring 16-index blocks share an XCD (workgroup w runs on XCD
    // w % 8), so that the two halves of a 128-byte line meet in one L2
    int blk = blockIdx.x;
    if constexpr (!XKC) {
        const int grp = blk >> 4, o = blk & 15;
        if ((grp + 1) * 16 <= (int)gridDim.x) blk = grp * 16 + 2 * (o & 7) + (o >> 3);
    }
    const int s0 = blk * 16;
    // X: op(A) (tall) or op(B)^T (wide), index = long-side index
    const float *xp = WIDE ? a.B : a.A;
    const int xld = WIDE ? a.ldb : a.lda;
    const float *yp = WIDE ? a.A : a.B;
    const int yld = WIDE ? a.lda : a.ldb;

    // the wave's K range: an NW-th, rounded up to whole groups
    const int kq4 = ((a.K + GK * NW - 1) / (GK * NW)) * GK;
    const int kb = wave * kq4, ke = min(a.K, kb + kq4);
    const int G = ke > kb ? (ke - kb + GK - 1) / GK : 0;

    SkinnyOperand<XKC, V> X;
    X.init(xp, L, xld, a.K, s0 + r, kq);
    SkinnyOperand<YKC, V> Y[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) Y[c].init(yp, S, yld, a.K, 16 * c + r, kq);

    f32x4 acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};

    // D groups in flight; group g lives in stage g % D.  Loads past the
    // wave's last group are issued with nothing in range (kend = kb), so
    // every iteration issues the same loads and the waits stay exact.
    float xv[D][V], yv[D][NC][V];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const int k0 = kb + GK * j, kend = j < G ? ke : kb;
        X.load(xv[j], k0, kq, kend);
#pragma unroll
        for (int c = 0; c < NC; ++c) Y[c].load(yv[j][c], k0, kq, kend);
    }
    for (int g0 = 0; g0 < G; g0 += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const int g = g0 + j;
            if (g < G) {
                if constexpr (!SPLIT) {
#pragma unroll
                    for (int s = 0; s < 4; ++s)
#pragma unroll
                        for (int c = 0; c < NC; ++c)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(xv[j][s], yv[j][c][s], acc[c], 0, 0, 0);
                } else {
                    // products of the pieces, smallest first, as the tile
                    // kernels take them (X plays A, Y plays B)
                    constexpr int NPROD = NP == 1 ? 1 : (NP == 2 ? 3 : 8);
                    constexpr int PA[8] = {NP == 3 ? 2 : 1, NP == 3 ? 1 : 0, NP == 3 ? 2 : 0, 1, 0, 1, 0, 0};
                    constexpr int PB[8] = {NP == 3 ? 1 : 0, NP == 3 ? 2 : 1, 0, 1, 2, 0, 1, 0};
                    bf16x8 xs[NP];
                    skinny_split<NP>(xv[j], xs);
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        bf16x8 ys[NP];
                        skinny_split<NP>(yv[j][c], ys);
#pragma unroll
                        for (int p = 0; p < NPROD; ++p) {
                            const int ia = NP == 1 ? 0 : PA[p], ib = NP == 1 ? 0 : PB[p];
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xs[ia], ys[ib], acc[c], 0, 0, 0);
                        }
                    }
                }
            }
            const int gn = g + D, k0 = kb + GK * gn, kend = gn < G ? ke : kb;
            X.load(xv[j], k0, kq, kend);
#pragma unroll
            for (int c = 0; c < NC; ++c) Y[c].load(yv[j][c], k0, kq, kend);
        }
    }

    // the waves' partial sums meet in LDS: D[i][j] of wave w at red[w][i][j]
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int t = 0; t < 4; ++t) red[wave][4 * q + t][16 * c + r] = acc[c][t];
    __syncthreads();

    // C = alpha * sum + beta * C (C not read when beta is 0).  Tall: D[i][j]
    // is C[s0 + i][j], consecutive threads along j (a row of C); wide: it is
    // C[j][s0 + i], consecutive threads along i.  In the split modes a sum
    // that is Inf or NaN is recomputed in FP32 (FMAs over k in order), so
    // special values give FP32's classes (the pieces of an Inf, or of a value
    // beyond the largest bf16, are not IEEE-faithful on their own).
    constexpr int E = 16 * 16 * NC;
#pragma unroll
    for (int e0 = 0; e0 < E; e0 += 64 * NW) {
        const int e = e0 + (int)threadIdx.x;
        if (e >= E) break;
        const int i = WIDE ? e % 16 : e / (16 * NC), j = WIDE ? e / 16 : e % (16 * NC);
        if (s0 + i >= L || j >= S) continue;
        float sum = 0.f;
#pragma unroll
        for (int w = 0; w < NW; ++w) sum += red[w][i][j];
        if constexpr (SPLIT) {
            if (__builtin_expect(!__builtin_isfinite(sum), 0)) {
                float v = 0.f;
                for (int k = 0; k < a.K; ++k)
                    v = __builtin_fmaf(SkinnyOperand<XKC, V>::at(xp, xld, s0 + i, k),
                                       SkinnyOperand<YKC, V>::at(yp, yld, j, k), v);
                sum = v;
            }
        }
        float *o = WIDE ? a.C + (long long)j * a.ldc + (s0 + i) : a.C + (long long)(s0 + i) * a.ldc + j;
        *o = a.beta != 0.f ? a.alpha * sum + a.beta * *o : a.alpha * sum;
    }
}

typedef hipError_t (*SkinnyLaunch)(const SkinnyArgs &a, hipStream_t s);

template <int PREC, int TA, int TB, int WIDE, int NC, int NW>
hipError_t launch_skinny(const SkinnyArgs &a, hipStream_t s)
{
    const int L = WIDE ? a.N : a.M;
    hipLaunchKernelGGL((k_gemm_skinny<PREC, TA, TB, WIDE, NC, NW>), dim3((unsigned)((L + 15) / 16)), dim3(64 * NW), 0,
                       s, a);
    return hipGetLastError();
}

// [precision FP32, BF16X3, BF16, BF16X6][ta][tb][wide][NC 1, 2, 4 -> 0, 1, 2]
// [NW 4, 8 -> 0, 1]; the split modes have the four-wave kernels with 1 or 2
// column blocks only.
struct SkinnyTable {
    SkinnyLaunch fn[4][2][2][2][3][2];
};
SkinnyTable &skinny_table();
void register_skinny_fp32(SkinnyTable &t);
void register_skinny_split(SkinnyTable &t);

}  // namespace ls
