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

// One GEMM (or one K slice of it) over a 1-D grid of tiles.  With ksplit > 1,
// blockIdx.y is the K slice and the tile goes to partial[slice] (M x N, fp32
// for the float modes, fp64 for FP64); a reduction kernel sums the slices in
// slice order, so results do not depend on scheduling.  Without split-K the
// tile is written as alpha * product + beta * C (C is not read when beta is 0).
template <class Cfg>
__global__ __launch_bounds__(Cfg::NT) void k_gemm_tiles(GemmDesc d, int tiles_n, int kper,
                                                        typename Cfg::T *partial, typename Cfg::T alpha,
                                                        typename Cfg::T beta)
{
    __shared__ __attribute__((aligned(16))) char lds[Cfg::LDS_BYTES];
    // Tiles in row-major order of blockIdx.x.  The dispatcher deals
    // workgroups to the 8 XCDs round-robin, so the resident tiles of one XCD
    // are every 8th of a band of rows: a compact set whose A and B panels
    // share that XCD's L2.  (Remapping each XCD to a contiguous run of
    // tiles made it sweep whole tile rows, every B panel through one L2,
    // and was up to 5% slower on large grids.)
    const int b = blockIdx.x;
    const int tm = b / tiles_n, tn = b - tm * tiles_n;
    const int m0 = tm * Cfg::BM, n0 = tn * Cfg::BN;
    const int kbeg = blockIdx.y * kper;
    const int kend = min(d.K, kbeg + kper);
    OneSrc src;
    if (gridDim.y == 1)
        src.job = tile_job<Cfg>(d, m0, n0, kbeg, kend, d.C, d.ldc);
    else
        src.job = tile_job<Cfg>(d, m0, n0, kbeg, kend, partial + (size_t)blockIdx.y * d.M * d.N, d.N);
    Tile<Cfg> t;
    t.scale = gridDim.y == 1 && !(alpha == typename Cfg::T(1) && beta == typename Cfg::T(0));
    t.alpha = alpha;
    t.beta = beta;
    t.stream(src, 0, 1, lds);
}

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

// C = alpha * (sum over slices of partial[s], in slice order) + beta * C, for
// the M x N partials and C with leading dimension ldc (C is not read when
// beta is 0).
template <typename T>
__global__ __launch_bounds__(256) void k_splitk_reduce(const T *__restrict__ partial, T *__restrict__ C,
                                                       long M, long N, long ldc, int nslices, T alpha, T beta)
{
    const long mn = M * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < mn; i += (long)gridDim.x * blockDim.x) {
        T s = partial[i];
        for (int j = 1; j < nslices; ++j) s += partial[(size_t)j * mn + i];
        const long r = i / N, c = i - r * N;
        T *o = C + r * ldc + c;
        *o = beta == T(0) ? alpha * s : alpha * s + beta * *o;
    }
}

// ---------------------------------------------------------------------------
// Launcher table (host side).
// ---------------------------------------------------------------------------
// alpha / beta are read as the operand type (float or double).
typedef hipError_t (*TilesLaunch)(const GemmDesc &d, int ksplit, int kper, void *partial, double alpha,
                                  double beta, hipStream_t s);
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
hipError_t launch_tiles(const GemmDesc &d, int ksplit, int kper, void *partial, double alpha, double beta,
                        hipStream_t s)
{
    typedef typename Cfg::T T;
    const int tiles_m = (d.M + Cfg::BM - 1) / Cfg::BM, tiles_n = (d.N + Cfg::BN - 1) / Cfg::BN;
    const int ntiles = tiles_m * tiles_n;
    dim3 grid(ntiles, ksplit);
    hipLaunchKernelGGL(k_gemm_tiles<Cfg>, grid, dim3(Cfg::NT), 0, s, d, tiles_n, kper, (T *)partial,
                       (T)alpha, (T)beta);
    return hipGetLastError();
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
    t.tiles[PREC][TA][TB][tile] = &launch_tiles<Cfg>;
    // the batched buckets use the square tiles
    if (BM == BN) {
        t.persist[PREC][TA][TB][tile] = &launch_persistent<Cfg>;
        t.persist_occ[PREC][TA][TB][tile] = &persistent_occupancy<Cfg>;
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
        typedef TileCfg<PREC, TA, TB, 16, 16> C16;
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
