// kernels.hpp -- __global__ entry points built on the tile GEMM, and the
// host-side launcher table each precision's translation unit fills in.
#pragma once
#include "gemm_core.hpp"

namespace ls {

// Tile shapes the planner can choose from (rows x cols of C per workgroup).
enum TileId : int { T32x32 = 0, T64x64, T128x128, T128x32, T128x64, T32x128, NTILE };
struct TileShape { int bm, bn; };
constexpr TileShape kTileShapes[NTILE] = {{32, 32}, {64, 64}, {128, 128}, {128, 32}, {128, 64}, {32, 128}};
constexpr int NPREC = 5;

// The K slices of one split-K tile summed in slice order into C (the
// workgroup that arrives last at the tile): C = alpha * sum + beta * C over
// the tile's tr x tc elements, C not read when beta is 0.  Each thread owns
// NG groups of 16 bytes of the tile and keeps the loads of a batch of SB
// slices for all of them in flight before adding (a round trip per slice
// batch, not per element; eight 16-byte loads in flight keep the kernel's
// register count where its tile loop puts it).
template <typename T, int BM, int BN, int NT>
__device__ __forceinline__ void splitk_sum(const T *__restrict__ slots, int ks, T *__restrict__ C, int ldc, int tr,
                                           int tc, T alpha, T beta)
{
    constexpr int TILE = BM * BN, V = 16 / sizeof(T), NG = TILE / V / NT;
    static_assert(NG * V * NT == TILE, "whole 16-byte groups per thread");
    constexpr int SB = NG >= 8 ? 1 : 8 / NG;  // slices per batch
    typedef T vec __attribute__((ext_vector_type(V)));
    vec acc[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) acc[g] = *(const vec *)(slots + (size_t)(threadIdx.x + g * NT) * V);
    for (int s0 = 1; s0 < ks; s0 += SB) {
        vec t[SB][NG];
#pragma unroll
        for (int j = 0; j < SB; ++j)
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (s0 + j < ks) t[j][g] = *(const vec *)(slots + (size_t)(s0 + j) * TILE + (threadIdx.x + g * NT) * V);
#pragma unroll
        for (int j = 0; j < SB; ++j)
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (s0 + j < ks) acc[g] += t[j][g];
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int e = (threadIdx.x + g * NT) * V, r = e / BN, c0 = e - r * BN;
        if (r >= tr) continue;
        T *o = C + (size_t)r * ldc + c0;
#pragma unroll
        for (int i = 0; i < V; ++i)
            if (c0 + i < tc) o[i] = beta == T(0) ? alpha * acc[g][i] : alpha * acc[g][i] + beta * o[i];
    }
}

// The XCD (accelerator complex die) the calling wave runs on: hardware
// register XCC_ID (id 20), bits 3:0.
__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15; }

// Workgroup linear ids of a split-K launch.  The dispatcher deals the
// workgroups of a 1-D launch to the 8 XCDs round-robin: workgroup L runs on
// XCD (L + o) % 8, where the start o depends on the hardware queue and its
// history (verified on the device before such a launch is used).  So
// workgroup L = ((t / 8) * ks + s) * 8 + t % 8 runs slice s of tile t on XCD
// (t + o) % 8 for every s: the slices of a tile share one L2.  Returns the
// tile, or -1 for the padding past the last tile.
__device__ __forceinline__ int splitk_tile(int L, int ks, int ntiles, int *slice)
{
    const int t = (L / (8 * ks)) * 8 + (L & 7);
    *slice = (L >> 3) % ks;
    return t < ntiles ? t : -1;
}

// After a slice has written its partial tile: publish it and count the
// arrival at tile t (counters[2t]: arrivals; counters[2t + 1]: the XCDs the
// slices ran on, one bit each); returns to every thread 0 unless this
// workgroup is the last slice, which then sees every slice's partial tile.
// With xcd_local (slices of a tile share an XCD, splitk_tile), a slice waits
// for its stores -- the L1 is write-through, so they are in the XCD's L2,
// where the last slice reads them -- and the last slice returns 2; should
// the slices have run on different XCDs after all, it returns 3 and the
// caller must not trust the partial tiles.  Otherwise every slice writes its
// L2 back and the last one invalidates its caches (agent-scope release and
// acquire) and returns 1.  The words return to zero for the next launch.
__device__ __forceinline__ int splitk_arrive(int *counters, int t, int ks, bool xcd_local, int *flag)
{
    if (xcd_local) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
    } else {
        __threadfence();
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int bit = 1 << xcc_id();
        atomicOr(&counters[2 * t + 1], bit);
        const bool last = atomicAdd(&counters[2 * t], 1) == ks - 1;
        int f = 0;
        if (last) {
            const int xcds = atomicOr(&counters[2 * t + 1], 0);
            f = !xcd_local ? 1 : xcds == bit ? 2 : 3;
            counters[2 * t] = 0;
            counters[2 * t + 1] = 0;
        }
        *flag = f;
    }
    __syncthreads();
    const int f = *flag;
    if (f == 1)
        __threadfence();
    else if (f == 2)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return f;
}

// The tr x tc tile of C at C (leading dimension ldc) set to NaN: the loud
// result of a split reduction whose partial tiles could not be trusted.
template <typename T, int NT>
__device__ __forceinline__ void nan_tile(T *C, int ldc, int tr, int tc)
{
    for (int e = threadIdx.x; e < tr * tc; e += NT) C[(size_t)(e / tc) * ldc + e % tc] = __builtin_nan("");
}

// One GEMM over a grid of tiles x K slices.  Without split-K (ks == 1)
// workgroup b is tile b in row-major order and writes alpha * product + beta
// * C (C is not read when beta is 0).  With it (the reference accumulates
// its K blocks in one pass, k_tag in Kernel/SMM_m8n12_new.h, SMM_NN_new.c:
// 438-459; here the slices run in parallel), the slices of a tile run on one
// XCD (splitk_tile), every slice writes its partial tile to its slot, and
// the last slice to arrive (splitk_arrive) sums the slots in slice order --
// deterministic, in the same launch -- and applies alpha / beta.  Counters
// take two ints per tile (splitk_arrive).
//
// Tiles in row-major order: the dispatcher deals workgroups to the 8 XCDs
// round-robin, so the resident tiles of one XCD are every 8th of a band of
// rows, a compact set whose A and B panels share that XCD's L2.  (Remapping
// each XCD to a contiguous run of tiles made it sweep whole tile rows, every
// B panel through one L2, and was up to 5% slower on large grids.)
template <class Cfg>
__global__ __launch_bounds__(Cfg::NT) void k_gemm_tiles(GemmDesc d, int tiles_n, int ntiles, int ks, int kper,
                                                        typename Cfg::T alpha, typename Cfg::T beta,
                                                        typename Cfg::T *slots, int *counters, int xcd_local)
{
    typedef typename Cfg::T T;
    __shared__ __attribute__((aligned(16))) char lds[Cfg::LDS_BYTES];
    __shared__ int flag;
    int slice = 0;
    const int b = ks == 1 ? (int)blockIdx.x : splitk_tile(blockIdx.x, ks, ntiles, &slice);
    if (b < 0) return;
    const int tm = b / tiles_n, tn = b - tm * tiles_n;
    const int m0 = tm * Cfg::BM, n0 = tn * Cfg::BN;
    const int kbeg = slice * kper;
    const int kend = min(d.K, kbeg + kper);
    T *slot = slots + ((size_t)b * ks + slice) * (Cfg::BM * Cfg::BN);
    OneSrc src;
    src.job = tile_job<Cfg>(d, m0, n0, kbeg, kend, d.C, d.ldc);
    if (ks > 1) {  // the partial tile goes to the slice's slot (leading dimension BN)
        src.job.C = slot;
        src.job.ldc = Cfg::BN;
    }
    Tile<Cfg> t;
    t.scale = ks == 1 && !(alpha == T(1) && beta == T(0));
    t.alpha = alpha;
    t.beta = beta;
    t.stream(src, 0, 1, lds);
    if (ks == 1) return;
    const int f = splitk_arrive(counters, b, ks, xcd_local != 0, &flag);
    if (f == 0) return;
    T *Ct = (T *)d.C + (size_t)m0 * d.ldc + n0;
    if (f == 3)
        nan_tile<T, Cfg::NT>(Ct, d.ldc, min(Cfg::BM, d.M - m0), min(Cfg::BN, d.N - n0));
    else
        splitk_sum<T, Cfg::BM, Cfg::BN, Cfg::NT>(slots + (size_t)b * ks * (Cfg::BM * Cfg::BN), ks, Ct, d.ldc,
                                                 min(Cfg::BM, d.M - m0), min(Cfg::BN, d.N - n0), alpha, beta);
}

// Records the XCD of each of n workgroups (see xcc_id).
__global__ void k_xcc_probe(int *out, int n);

// Persistent executor: as many workgroups as can be resident stream the job
// list, workgroup b taking jobs b, b + grid, ...  Two resident workgroups per
// CU for the 128 x 128 tiles, which keeps two job streams' loads in flight.
template <class Cfg>
__global__ __attribute__((amdgpu_flat_work_group_size(Cfg::NT, Cfg::NT),
                          amdgpu_waves_per_eu(Cfg::WAVES_PER_EU, 8))) void
k_gemm_persistent(const GemmDesc *__restrict__ jobs, int njobs)
{
    __shared__ __attribute__((aligned(16))) char lds[Cfg::LDS_BYTES];
    ArraySrc src{jobs, njobs};
    Tile<Cfg> t;
    t.stream(src, blockIdx.x, gridDim.x, lds);
}

// ---------------------------------------------------------------------------
// Launcher table (host side).
// ---------------------------------------------------------------------------
// alpha / beta are read as the operand type (float or double).  With
// ksplit > 1, slots holds ksplit partial tiles (BM x BN elements of the
// operand type) per tile and counters one zeroed int per tile.
typedef hipError_t (*TilesLaunch)(const GemmDesc &d, int ksplit, int kper, void *slots, int *counters,
                                  double alpha, double beta, bool xcd_local, hipStream_t s);
typedef hipError_t (*PersistLaunch)(const GemmDesc *jobs, int njobs, int grid, hipStream_t s);
typedef int (*PersistOccupancy)();  // resident workgroups per CU, 0 on error

struct LaunchTable {
    TilesLaunch tiles[NPREC][2][2][NTILE];
    PersistLaunch persist[NPREC][2][2][NTILE];
    PersistOccupancy persist_occ[NPREC][2][2][NTILE];
    // FP64 batched plans only: one-wave 16 x 16 tiles for the GEMMs with
    // M, N <= 16 (a 32 x 32 tile's four waves would leave three idle)
    PersistLaunch persist16[2][2];
    PersistOccupancy persist16_occ[2][2];
};
LaunchTable &launch_table();

template <class Cfg>
hipError_t launch_tiles(const GemmDesc &d, int ksplit, int kper, void *slots, int *counters, double alpha,
                        double beta, bool xcd_local, hipStream_t s)
{
    typedef typename Cfg::T T;
    const int tiles_m = (d.M + Cfg::BM - 1) / Cfg::BM, tiles_n = (d.N + Cfg::BN - 1) / Cfg::BN;
    const int ntiles = tiles_m * tiles_n;
    // split-K: tiles padded to a multiple of 8, times the slices (splitk_tile)
    const long grid = ksplit == 1 ? ntiles : (long)((ntiles + 7) / 8) * 8 * ksplit;
    hipLaunchKernelGGL(k_gemm_tiles<Cfg>, dim3((unsigned)grid), dim3(Cfg::NT), 0, s, d, tiles_n, ntiles, ksplit,
                       kper, (T)alpha, (T)beta, (T *)slots, counters, (int)xcd_local);
    return hipGetLastError();
}

// Bytes of split-K partial-tile slots a plan needs: ksplit tiles of BM x BN
// elements per tile of C.
inline size_t splitk_slot_bytes(int tile, long M, long N, int ksplit, int esize)
{
    if (ksplit <= 1) return 0;
    const TileShape ts = kTileShapes[tile];
    const size_t ntiles = (size_t)((M + ts.bm - 1) / ts.bm) * (size_t)((N + ts.bn - 1) / ts.bn);
    return ntiles * (size_t)ksplit * ts.bm * ts.bn * esize;
}

template <class Cfg>
hipError_t launch_persistent(const GemmDesc *jobs, int njobs, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_gemm_persistent<Cfg>, dim3(grid), dim3(Cfg::NT), 0, s, jobs, njobs);
    return hipGetLastError();
}

template <class Cfg>
int persistent_occupancy()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_gemm_persistent<Cfg>, Cfg::NT, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

template <int PREC, int TA, int TB, int BM, int BN>
void register_one(LaunchTable &t, int tile)
{
    typedef TileCfg<PREC, TA, TB, BM, BN> Cfg;
    typedef TileCfg<PREC, TA, TB, BM, BN, true> PCfg;
    t.tiles[PREC][TA][TB][tile] = &launch_tiles<Cfg>;
    // the batched buckets use the square tiles
    if (BM == BN) {
        t.persist[PREC][TA][TB][tile] = &launch_persistent<PCfg>;
        t.persist_occ[PREC][TA][TB][tile] = &persistent_occupancy<PCfg>;
    }
}

template <int PREC, int TA, int TB>
void register_trans(LaunchTable &t)
{
    register_one<PREC, TA, TB, 32, 32>(t, T32x32);
    register_one<PREC, TA, TB, 64, 64>(t, T64x64);
    register_one<PREC, TA, TB, 128, 128>(t, T128x128);
    register_one<PREC, TA, TB, 128, 32>(t, T128x32);
    register_one<PREC, TA, TB, 128, 64>(t, T128x64);
    register_one<PREC, TA, TB, 32, 128>(t, T32x128);
    if constexpr (PREC == P_FP64) {
        typedef TileCfg<PREC, TA, TB, 16, 16, true> C16;
        t.persist16[TA][TB] = &launch_persistent<C16>;
        t.persist16_occ[TA][TB] = &persistent_occupancy<C16>;
    }
}

// Defined in one translation unit per precision and transa.
void register_fp32_ta0(LaunchTable &t);
void register_fp32_ta1(LaunchTable &t);
void register_bf16x3_ta0(LaunchTable &t);
void register_bf16x3_ta1(LaunchTable &t);
void register_bf16_ta0(LaunchTable &t);
void register_bf16_ta1(LaunchTable &t);
void register_bf16x6_ta0(LaunchTable &t);
void register_bf16x6_ta1(LaunchTable &t);
void register_fp64_ta0(LaunchTable &t);
void register_fp64_ta1(LaunchTable &t);

}  // namespace ls
