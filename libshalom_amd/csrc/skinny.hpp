// skinny.hpp -- one FP32 GEMM with a short side: a tall C (N <= 64, M large)
// or a wide one (M <= 64, N large), the irregular shapes of sgemm_mp
// (SMM_thread.c:75-114 splits their long side over threads).
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
#pragma once
#include "gemm_core.hpp"

namespace ls {

struct SkinnyArgs {
    const float *A, *B;
    float *C;
    int M, N, K, lda, ldb, ldc;
    float alpha, beta;
};

// Groups of 16 k in flight per wave: more for the narrower small sides,
// whose groups are fewer bytes.
template <int NC>
constexpr int skinny_depth() { return NC == 1 ? 8 : NC == 2 ? 6 : 4; }

// One operand's loads for a group: KC, one 16-byte load per lane; IC, four
// 4-byte loads.  Elements past the wave's K range or past the operand's
// index count read 0.
template <bool KC>
struct SkinnyOperand {
    __amdgpu_buffer_rsrc_t rs;
    int base[KC ? 1 : 4];  // byte offset of the lane's k0 = 0 elements
    int kstep;             // bytes per k
    bool idx_ok;           // the lane's index is inside the operand

    // index i of this lane, k offset kq = 4q of the lane within a group
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
            for (int s = 0; s < 4; ++s) base[s] = ((kq + s) * ld + i) * 4;
            kstep = ld * 4;
        }
    }

    // Elements k0 + kq + s (s = 0..3) of the lane's index; k >= kend reads
    // 0 (the next wave's range, or past K: in range for a KC row).
    __device__ __forceinline__ void load(float (&v)[4], int k0, int kq, int kend) const
    {
        const int koff = k0 * kstep;
        if constexpr (KC) {
            float t[4];
            buf_load16<float>(rs, oob_unless(k0 < kend, base[0] + koff), t);
#pragma unroll
            for (int s = 0; s < 4; ++s) v[s] = (k0 + kq + s < kend) ? t[s] : 0.f;
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s)
                v[s] = buf_load1<float>(rs, oob_unless(idx_ok & (k0 + kq + s < kend), base[s] + koff));
        }
    }
};

template <int TA, int TB, int WIDE, int NC, int NW>
__global__ __launch_bounds__(64 * NW) void k_gemm_skinny(SkinnyArgs a)
{
    constexpr int D = skinny_depth<NC>();
    // X: k-contiguous when op(A) is A (tall) or op(B)^T is B stored N x K (wide)
    constexpr bool XKC = WIDE ? TB != 0 : TA == 0;
    constexpr bool YKC = WIDE ? TA == 0 : TB != 0;
    __shared__ __attribute__((aligned(16))) float red[NW][16][16 * NC];

    const int lane = threadIdx.x & 63, wave = uniform((int)(threadIdx.x >> 6));
    const int r = lane & 15, q = lane >> 4, kq = 4 * q;
    const int L = WIDE ? a.N : a.M;  // long side
    const int S = WIDE ? a.M : a.N;  // short side
    const int s0 = blockIdx.x * 16;
    // X: op(A) (tall) or op(B)^T (wide), index = long-side index
    const float *xp = WIDE ? a.B : a.A;
    const int xld = WIDE ? a.ldb : a.lda;
    const float *yp = WIDE ? a.A : a.B;
    const int yld = WIDE ? a.lda : a.ldb;

    // the wave's K range: an NW-th, rounded up to whole groups of 16
    const int kq4 = ((a.K + 16 * NW - 1) / (16 * NW)) * 16;
    const int kb = wave * kq4, ke = min(a.K, kb + kq4);
    const int G = ke > kb ? (ke - kb + 15) / 16 : 0;

    SkinnyOperand<XKC> X;
    X.init(xp, L, xld, a.K, s0 + r, kq);
    SkinnyOperand<YKC> Y[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) Y[c].init(yp, S, yld, a.K, 16 * c + r, kq);

    f32x4 acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};

    // D groups in flight; group g lives in stage g % D.  Loads past the
    // wave's last group are issued with nothing in range (kend = kb), so
    // every iteration issues the same loads and the waits stay exact.
    float xv[D][4], yv[D][NC][4];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const int k0 = kb + 16 * j, kend = j < G ? ke : kb;
        X.load(xv[j], k0, kq, kend);
#pragma unroll
        for (int c = 0; c < NC; ++c) Y[c].load(yv[j][c], k0, kq, kend);
    }
    for (int g0 = 0; g0 < G; g0 += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const int g = g0 + j;
            if (g < G) {
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int c = 0; c < NC; ++c)
                        acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(xv[j][s], yv[j][c][s], acc[c], 0, 0, 0);
            }
            const int gn = g + D, k0 = kb + 16 * gn, kend = gn < G ? ke : kb;
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
    // C[j][s0 + i], consecutive threads along i.
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
        float *o = WIDE ? a.C + (long long)j * a.ldc + (s0 + i) : a.C + (long long)(s0 + i) * a.ldc + j;
        *o = a.beta != 0.f ? a.alpha * sum + a.beta * *o : a.alpha * sum;
    }
}

typedef hipError_t (*SkinnyLaunch)(const SkinnyArgs &a, hipStream_t s);

template <int TA, int TB, int WIDE, int NC, int NW>
hipError_t launch_skinny(const SkinnyArgs &a, hipStream_t s)
{
    const int L = WIDE ? a.N : a.M;
    hipLaunchKernelGGL((k_gemm_skinny<TA, TB, WIDE, NC, NW>), dim3((unsigned)((L + 15) / 16)), dim3(64 * NW), 0, s,
                       a);
    return hipGetLastError();
}

// [ta][tb][wide][NC 1, 2, 4 -> 0, 1, 2][NW 4, 8 -> 0, 1]
struct SkinnyTable {
    SkinnyLaunch fn[2][2][2][3][2];
};
SkinnyTable &skinny_table();

}  // namespace ls
