// backward.hpp -- the fused backward launch of C = op(A) op(B):
//   op(A)' = dC op(B)^T   and   op(B)' = op(A)^T dC
// for one GEMM or a strided batch of them, in ONE kernel launch.
//
// Each of the two gradients is a GEMM problem of its own ("X" = dA, "Y" =
// dB): an output of R x Cc, a left operand R x L and a right operand L x Cc
// (each stored plain or transposed), summed over a reduction of length L per
// image -- and over the images too when the operand whose gradient it is is
// shared by the batch (stride 0, e.g. the weights of a convolution).  Every
// workgroup of the launch owns one unit: one output tile of X or of Y and one
// slice of its reduction.  A slice is a run of segments (an image, or part of
// an image's reduction) chained through the same accumulators.  When a tile's
// reduction is cut into several slices, each slice writes its partial tile to
// a slot, and the last slice to arrive (per-tile arrival counter) sums the
// slots in slice order and writes the tile -- deterministic, inside the same
// launch, and the counter is reset to zero for the next launch.
//
// The X and Y units that read the same dC block run in the same launch, but
// a launch's workgroups spread over the eight XCDs, each with its own L2, so
// a block read by tiles on several XCDs is fetched several times: measured
// HBM traffic of the stored-im2col convolution backward is 1.8-2.9x its
// algorithmic bytes (operands read once, dW and dX written once), the excess
// being operand re-reads and the slice slots.
#pragma once
#include "kernels.hpp"

namespace ls {

struct BwdProblem {
    const void *L, *Rt;  // left / right operand of image 0
    void *O;             // output of image 0 (the only one when shared)
    long long sL, sR, sO;  // per-image strides in elements; sO == 0: one output summed over images
    int lda, ldb, ldo;
    int R, Cc, red;      // output rows, columns; reduction length per image
    int nimg;
    int tiles_c, ntiles;  // tile grid of one output
    int nslices;          // reduction slices per output tile
    int spi;              // > 0: slices per image (a slice lies inside one image)
    int gimg;             // spi == 0: whole images per slice
    int slen;             // spi > 0: reduction length of a slice (multiple of kalign)
    int kalign;           // slice lengths are multiples of this (the K chunk, and a
                          // convolution's R*S so that a slice starts on a channel)
    int units;            // units of work of this problem
    int grid;             // workgroups of this problem: units, and with slices the
                          // tiles padded to a multiple of 8 (unit_of)
    int xcd;              // slices of a tile share an XCD (see splitk_arrive)
    float *slots;         // nslices > 1: partial tiles, [tile][slice][BM * BN]
    int *counters;        // nslices > 1: two ints per tile (splitk_arrive; zero between launches)
    // A convolution output with the batch folded into its columns (fold_total
    // > 0): column n of the R x Cc output is pixel n % ldo of image n / ldo,
    // at n + (n / ldo) * (R - 1) * ldo; fold_total = all output elements.
    int fold_total;
    unsigned fold_mpq;    // ceil(2^32 / ldo)
};

// The segments of one unit's slice as a job stream (see Tile::stream).  The
// problem's fields are copied in (scalar registers): a pointer to the kernel
// argument would send it through private memory.  With IM2COL_ = kIm2col
// the right operand of every image is the im2col matrix of the convolution
// `geom` over the image Rt + img * sR: the job's B is that image at the
// slice's first channel and its ldb the tile's first output pixel.  With
// kIm2colT the right operand is stored transposed and is that im2col
// matrix (rows = taps from c0, reduction = output pixels from rho0); the
// job's B is the image.
template <class Cfg, int IM2COL_ = kPlainB>
struct SliceSrc {
    static constexpr int IM2COL = IM2COL_;
    static constexpr int REPAIR = kRepairByCaller;
    const void *L, *Rt;
    long long sL, sR;
    int lda, ldb;
    int R, Cc;
    int img0, nseg;   // first image and number of segments
    int rho0, rho1;   // reduction range of every segment
    int r0, c0;       // output tile origin
    long long boff;   // offset of the segment's first B element in its image
    void *out;        // where the tile goes: the output, or the unit's slot
    int ldo;
    ConvGeom geom;
    FoldedCols fcols;   // IM2COL_ == kIm2col: where the tile's columns land

    __device__ __forceinline__ int count() const { return nseg; }
    __device__ __forceinline__ auto cols() const
    {
        if constexpr (IM2COL == kIm2col)
            return fcols;
        else
            return PlainCols{};
    }
    __device__ __forceinline__ GemmDesc get(int j) const
    {
        typedef typename Cfg::T T;
        const int img = img0 + j;
        const T *Li = (const T *)L + (long long)img * sL;
        const T *Ri = (const T *)Rt + (long long)img * sR;
        GemmDesc d;
        d.A = Li + (Cfg::TA ? (long long)rho0 * lda + r0 : (long long)r0 * lda + rho0);
        d.B = Ri + boff;
        d.C = out;
        d.M = min(Cfg::BM, R - r0);
        d.N = min(Cfg::BN, Cc - c0);
        d.K = rho1 - rho0;
        d.lda = lda;
        d.ldb = IM2COL == kIm2col ? c0 : ldb;
        d.ldc = ldo;
        d.flags = j + 1 < nseg ? kChain : 0;
        return d;
    }
};

// Sum the slots of tile tg in slice order into the output (the last slice
// to arrive, see splitk_arrive).
template <int BM, int BN, int NT>
__device__ __forceinline__ void bwd_fixup(const BwdProblem &P, int tg, int img_out, int r0, int c0, int *flag)
{
    const int f = splitk_arrive(P.counters, tg, P.nslices, P.xcd != 0, flag);
    if (f == 0) return;
    const int tr = min(BM, P.R - r0), tcn = min(BN, P.Cc - c0);
    const float *base = P.slots + (size_t)tg * P.nslices * (BM * BN);
    float *O = (float *)P.O + (long long)img_out * P.sO;
    for (int e = threadIdx.x * 4; e < tr * BN; e += NT * 4) {
        const int r = e / BN, c = e % BN;
        if (c >= tcn) continue;
        // slices summed in slice order; eight loads in flight at a time (a
        // load-then-add loop put one memory round trip per slice in series)
        f32x4 acc = *(const f32x4 *)(base + e);
        if (f == 3) acc = f32x4{NAN, NAN, NAN, NAN};  // partial tiles not to be trusted
        int s = f == 3 ? P.nslices : 1;
        for (; s + 8 <= P.nslices; s += 8) {
            f32x4 t[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) t[j] = *(const f32x4 *)(base + (size_t)(s + j) * (BM * BN) + e);
#pragma unroll
            for (int j = 0; j < 8; ++j) acc += t[j];
        }
        for (; s < P.nslices; ++s) acc += *(const f32x4 *)(base + (size_t)s * (BM * BN) + e);
        float *o = O + (long long)(r0 + r) * P.ldo;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (c + q >= tcn) continue;
            const int n = c0 + c + q;  // column of the output
            const long long col = P.fold_total > 0 ? n + (long long)udiv(n, P.fold_mpq) * (P.R - 1) * P.ldo : n;
            o[col] = acc[q];
        }
    }
}

// Unit of workgroup u of a problem: tile group tg (output tile, and image
// for unshared outputs) and slice s.  With slices, workgroup u runs slice
// (u / 8) % nslices of tile group (u / (8 * nslices)) * 8 + u % 8, so the
// slices of a tile share XCD tg % 8 (workgroup u runs on XCD u % 8; the
// launch starts every problem at a multiple of 8).  tg < 0 marks the
// padding past the last tile group.
struct Unit {
    int tg, s;  // tg < 0: padding
};
__device__ __forceinline__ Unit unit_of(int nslices, int units, int u)
{
    if (nslices == 1) return Unit{u < units ? u : -1, 0};
    const int tg = (u / (8 * nslices)) * 8 + (u & 7);
    return Unit{tg < units / nslices ? tg : -1, (u >> 3) % nslices};
}

// The job source of workgroup u of problem P (the tile group and slice of
// unit_of), or tg < 0 for padding past the last tile group.
template <class Cfg, int IM2COL>
struct UnitSrc {
    SliceSrc<Cfg, IM2COL> src;
    int tg, img_out;
};
template <class Cfg, int IM2COL>
__device__ __forceinline__ UnitSrc<Cfg, IM2COL> unit_src(const BwdProblem &P, int u, const ConvGeom &geom)
{
    UnitSrc<Cfg, IM2COL> us;
    const Unit un = unit_of(P.nslices, P.units, u);
    us.tg = un.tg;
    if (un.tg < 0) return us;
    const int tg = un.tg, s = un.s;
    const bool shared = P.sO == 0;
    const int img_out = shared ? 0 : tg / P.ntiles;
    us.img_out = img_out;
    const int t = shared ? tg : tg - img_out * P.ntiles;
    const int tr = t / P.tiles_c, tc = t - tr * P.tiles_c;
    SliceSrc<Cfg, IM2COL> &src = us.src;
    src.geom = geom;
    src.L = P.L;
    src.Rt = P.Rt;
    src.sL = P.sL;
    src.sR = P.sR;
    src.lda = P.lda;
    src.ldb = P.ldb;
    src.R = P.R;
    src.Cc = P.Cc;
    src.r0 = tr * Cfg::BM;
    src.c0 = tc * Cfg::BN;
    if (!shared) {
        src.img0 = img_out;
        src.nseg = 1;
        src.rho0 = s * P.slen;
        src.rho1 = min(P.red, src.rho0 + P.slen);
    } else if (P.spi > 0) {
        const int img = s / P.spi, sub = s - img * P.spi;
        src.img0 = img;
        src.nseg = 1;
        src.rho0 = sub * P.slen;
        src.rho1 = min(P.red, src.rho0 + P.slen);
    } else {
        src.img0 = s * P.gimg;
        src.nseg = min(P.gimg, P.nimg - src.img0);
        src.rho0 = 0;
        src.rho1 = P.red;
    }
    // a convolution slice starts on a channel (slices are multiples of R*S)
    if constexpr (IM2COL == kIm2col)
        src.boff = (long long)(src.rho0 / geom.RS) * geom.H * geom.W;
    else if constexpr (IM2COL == kIm2colT || IM2COL == kIm2colTR)
        src.boff = 0;
    else
        src.boff = Cfg::TB ? (long long)src.c0 * P.ldb + src.rho0 : (long long)src.rho0 * P.ldb + src.c0;
    // the tile's columns: folded convolution output written directly, or
    // the tile's own columns (FoldedCols with no skip and the plain extent)
    const bool folded = IM2COL == kIm2col && P.fold_total > 0 && P.nslices == 1;
    src.fcols.n0 = folded ? src.c0 : 0;
    src.fcols.skip = folded ? (P.R - 1) * P.ldo : 0;
    src.fcols.ext = folded ? P.fold_total - src.r0 * P.ldo
                           : (min(Cfg::BM, P.R - src.r0) - 1) * (P.nslices == 1 ? P.ldo : Cfg::BN) +
                                 min(Cfg::BN, P.Cc - src.c0);
    src.fcols.mPQ = P.fold_mpq;
    if (P.nslices == 1) {
        src.out = (float *)P.O + (long long)img_out * P.sO + (long long)src.r0 * P.ldo + (folded ? 0 : src.c0);
        src.ldo = P.ldo;
    } else {
        src.out = P.slots + ((size_t)tg * P.nslices + s) * (Cfg::BM * Cfg::BN);
        src.ldo = Cfg::BN;
    }
    return us;
}

// The split modes' FP32 recompute of the Inf / NaN elements of unit u's
// tile once its stream has stored it (Tile::rescan; a rare path).  The
// problem and the geometry are read again from the kernel's argument
// segment (p_off, g_off: their byte offsets there) through a pointer the
// compiler cannot see through: reusing the registers that hold them would
// keep those live through the stream loop and raise the kernels' register
// counts.
typedef const __attribute__((address_space(4))) char *KernArgs;

template <class Cfg, int IM2COL>
__device__ __forceinline__ void bwd_rescan(int p_off, int g_off, int u)
{
#if __HIP_DEVICE_COMPILE__  // (the host pass only parses device code)
    typedef const __attribute__((address_space(4))) BwdProblem *PP;
    typedef const __attribute__((address_space(4))) ConvGeom *GP;
    KernArgs args = (KernArgs)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(args));
    const BwdProblem P = *(PP)(args + p_off);
    ConvGeom g{};
    if constexpr (IM2COL != kPlainB) g = *(GP)(args + g_off);
    const UnitSrc<Cfg, IM2COL> us = unit_src<Cfg, IM2COL>(P, u, g);
    Tile<Cfg>::rescan(us.src, us.src.nseg - 1, us.src.nseg - 1, 1);
#endif
}

// Byte offsets of the arguments of the kernels below in their argument
// segment (laid out in order, each at its natural alignment).
struct ArgsXY {
    BwdProblem X, Y;
};
struct ArgsPG {
    BwdProblem P;
    ConvGeom g;
};
struct ArgsXYGG {
    BwdProblem X, Y;
    ConvGeom gx, gy;
};

// Unit u of problem P, the kernel argument at byte offset p_off (geom: the
// one at g_off).
template <class Cfg, int IM2COL = kPlainB>
__device__ __forceinline__ void bwd_unit(const BwdProblem &P, int p_off, int u, char *lds, int *flag,
                                         const ConvGeom &geom = {}, int g_off = 0)
{
    const UnitSrc<Cfg, IM2COL> us = unit_src<Cfg, IM2COL>(P, u, geom);
    if (us.tg < 0) return;
    const SliceSrc<Cfg, IM2COL> &src = us.src;
    Tile<Cfg> tile;
    if (tile.stream(src, 0, 1, lds)) bwd_rescan<Cfg, IM2COL>(p_off, g_off, u);
    if (P.nslices > 1) bwd_fixup<Cfg::BM, Cfg::BN, Cfg::NT>(P, us.tg, us.img_out, src.r0, src.c0, flag);
}

template <class CX, class CY>
struct BwdLds {
    static constexpr int BYTES = CX::LDS_BYTES > CY::LDS_BYTES ? CX::LDS_BYTES : CY::LDS_BYTES;
};

// Which problem workgroup u of a two-problem launch works on, and its unit
// there.  The problem with the longer units comes first (yfirst: Y), so that
// those start in the first round rather than behind the other's; the second
// starts on a multiple of 8, so that a unit's index keeps its XCD.
struct TwoUnits {
    bool x;
    int u;
};
__device__ __forceinline__ TwoUnits two_units(const BwdProblem &X, const BwdProblem &Y, int yfirst, int u)
{
    const int xg = (X.grid + 7) & ~7, yg = (Y.grid + 7) & ~7;
    if (yfirst) return u >= yg ? TwoUnits{true, u - yg} : TwoUnits{false, u};
    return u < xg ? TwoUnits{true, u} : TwoUnits{false, u - xg};
}

// Reduction steps of a problem's longest unit (0: none); a two-problem
// launch puts the problem with the longer ones first.
inline long longest_unit(const BwdProblem &P)
{
    return P.units <= 0 ? 0 : P.spi > 0 ? (long)P.slen : (long)P.gimg * P.red;
}
inline int launch_grid(const BwdProblem &X, const BwdProblem &Y, int yfirst)
{
    return yfirst ? ((Y.grid + 7) & ~7) + X.grid : ((X.grid + 7) & ~7) + Y.grid;
}

// dA units and dB units in one launch (two_units).
template <class CX, class CY>
__global__ __attribute__((amdgpu_flat_work_group_size(CX::NT, CX::NT))) void
k_gemm_backward(BwdProblem X, BwdProblem Y, int yfirst)
{
    __shared__ __attribute__((aligned(16))) char lds[BwdLds<CX, CY>::BYTES];
    __shared__ int flag;
    const TwoUnits w = two_units(X, Y, yfirst, blockIdx.x);
    if (w.x)
        bwd_unit<CX>(X, 0, w.u, lds, &flag);
    else
        bwd_unit<CY>(Y, offsetof(ArgsXY, Y), w.u, lds, &flag);
}

// One problem on its own: a strided batch of GEMMs C_i = op(A_i) op(B_i)
// (a stride of 0 shares an operand; a C stride of 0 sums over the batch).
template <class Cfg>
__global__ __attribute__((amdgpu_flat_work_group_size(Cfg::NT, Cfg::NT),
                          amdgpu_waves_per_eu(Cfg::BM * Cfg::BN >= 128 * 128 ? 2 : 1, 8))) void
k_gemm_strided(BwdProblem P)
{
    __shared__ __attribute__((aligned(16))) char lds[Cfg::LDS_BYTES];
    __shared__ int flag;
    bwd_unit<Cfg>(P, 0, blockIdx.x, lds, &flag);
}

// Convolution forward as an implicit GEMM: image i's output (Cout x P*Q) is
// the filter (Cout x C*R*S) times the im2col matrix of image i, whose
// elements the B-operand loads gather from the image directly.
template <class Cfg>
__global__ __attribute__((amdgpu_flat_work_group_size(Cfg::NT, Cfg::NT),
                          amdgpu_waves_per_eu(Cfg::BM * Cfg::BN >= 128 * 128 ? 2 : 1, 8))) void
k_conv_fwd(BwdProblem P, ConvGeom g)
{
    __shared__ __attribute__((aligned(16))) char lds[Cfg::LDS_BYTES];
    __shared__ int flag;
    bwd_unit<Cfg, kIm2col>(P, 0, blockIdx.x, lds, &flag, g, offsetof(ArgsPG, g));
}

// Convolution backward as two implicit GEMMs in ONE launch.  X units: the
// input gradient, itself a convolution of dY (geometry gx; fractionally
// strided for a strided convolution) with the filter flipped and transposed
// (C x Cout*R*S, in the problem's left operand).  Y units: the filter gradient
// dW (Cout x C*R*S) = sum over images of dY_i (Cout x P*Q) times the
// transposed im2col matrix of X_i, gathered from the image (geometry gy),
// the sum over images and pixels cut into slices as for a shared operand.
template <class CX, class CY, int IMY>
__global__ __attribute__((amdgpu_flat_work_group_size(CX::NT, CX::NT))) void
k_conv_backward(BwdProblem X, BwdProblem Y, ConvGeom gx, ConvGeom gy, int yfirst)
{
    __shared__ __attribute__((aligned(16))) char lds[BwdLds<CX, CY>::BYTES];
    __shared__ int flag;
    const TwoUnits w = two_units(X, Y, yfirst, blockIdx.x);
    if (w.x)
        bwd_unit<CX, kIm2col>(X, 0, w.u, lds, &flag, gx, offsetof(ArgsXYGG, gx));
    else
        bwd_unit<CY, IMY>(Y, offsetof(ArgsXY, Y), w.u, lds, &flag, gy, offsetof(ArgsXYGG, gy));
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
typedef hipError_t (*BwdLaunch)(const BwdProblem &X, const BwdProblem &Y, hipStream_t s);
typedef hipError_t (*StridedLaunch)(const BwdProblem &P, hipStream_t s);
typedef hipError_t (*ConvLaunch)(const BwdProblem &P, const ConvGeom &g, hipStream_t s);
typedef hipError_t (*ConvBwdLaunch)(const BwdProblem &X, const BwdProblem &Y, const ConvGeom &gx,
                                    const ConvGeom &gy, hipStream_t s);

// RUNS: dW's gather in 16-byte runs (stride 1 along W, Q a multiple of 4).
template <int PREC, int BT, bool RUNS>
hipError_t launch_conv_backward(const BwdProblem &X, const BwdProblem &Y, const ConvGeom &gx, const ConvGeom &gy,
                                hipStream_t s)
{
    typedef TileCfg<PREC, 0, 0, BT, BT> CX;  // dX: flipped filter x im2col(dY)
    typedef TileCfg<PREC, 0, 1, BT, BT> CY;  // dW: dY x im2col(X)^T
    static_assert(CX::NT == CY::NT, "one workgroup size for both problems");
    if (X.units + Y.units <= 0) return hipSuccess;
    const int yfirst = longest_unit(Y) > longest_unit(X);
    hipLaunchKernelGGL((k_conv_backward<CX, CY, RUNS ? kIm2colTR : kIm2colT>), dim3(launch_grid(X, Y, yfirst)),
                       dim3(CX::NT), 0, s, X, Y, gx, gy, yfirst);
    return hipGetLastError();
}

template <int PREC, int BT>
hipError_t launch_conv_fwd(const BwdProblem &P, const ConvGeom &g, hipStream_t s)
{
    typedef TileCfg<PREC, 0, 0, BT, BT> Cfg;
    if (P.units <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_conv_fwd<Cfg>, dim3(P.grid), dim3(Cfg::NT), 0, s, P, g);
    return hipGetLastError();
}

template <int PREC, int TA, int TB, int BT>
hipError_t launch_strided(const BwdProblem &P, hipStream_t s)
{
    typedef TileCfg<PREC, TA, TB, BT, BT> Cfg;
    if (P.units <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gemm_strided<Cfg>, dim3(P.grid), dim3(Cfg::NT), 0, s, P);
    return hipGetLastError();
}

// (transa, transb) of the forward product -> (TA, TB) of the dA and dB
// problems (left operand transposed / right operand transposed).
constexpr int kBwdTA[2][2][2] = {{{0, 1}, {0, 1}}, {{0, 0}, {1, 1}}};  // [ta][tb][X/Y]: left transposed
constexpr int kBwdTB[2][2][2] = {{{1, 0}, {0, 0}}, {{1, 0}, {1, 1}}};  // [ta][tb][X/Y]: right transposed

template <int PREC, int TA, int TB, int BT>
hipError_t launch_backward(const BwdProblem &X, const BwdProblem &Y, hipStream_t s)
{
    typedef TileCfg<PREC, kBwdTA[TA][TB][0], kBwdTB[TA][TB][0], BT, BT> CX;
    typedef TileCfg<PREC, kBwdTA[TA][TB][1], kBwdTB[TA][TB][1], BT, BT> CY;
    if (X.units + Y.units <= 0) return hipSuccess;
    static_assert(CX::NT == CY::NT, "one workgroup size for both problems");
    const int yfirst = longest_unit(Y) > longest_unit(X);
    hipLaunchKernelGGL((k_gemm_backward<CX, CY>), dim3(launch_grid(X, Y, yfirst)), dim3(CX::NT), 0, s, X, Y, yfirst);
    return hipGetLastError();
}

constexpr int NBWD_TILE = 2;  // 0: 64 x 64, 1: 128 x 128
struct BwdTable {
    BwdLaunch fn[NPREC][2][2][NBWD_TILE];
    StridedLaunch strided[NPREC][2][2][NBWD_TILE];
    ConvLaunch conv_fwd[NPREC][NBWD_TILE];
    ConvBwdLaunch conv_bwd[NPREC][NBWD_TILE][2];  // [.][.][dW in 16-byte runs]
};
BwdTable &bwd_table();

template <int PREC>
void register_backward(BwdTable &t)
{
    t.fn[PREC][0][0][0] = &launch_backward<PREC, 0, 0, 64>;
    t.fn[PREC][0][1][0] = &launch_backward<PREC, 0, 1, 64>;
    t.fn[PREC][1][0][0] = &launch_backward<PREC, 1, 0, 64>;
    t.fn[PREC][1][1][0] = &launch_backward<PREC, 1, 1, 64>;
    t.fn[PREC][0][0][1] = &launch_backward<PREC, 0, 0, 128>;
    t.fn[PREC][0][1][1] = &launch_backward<PREC, 0, 1, 128>;
    t.fn[PREC][1][0][1] = &launch_backward<PREC, 1, 0, 128>;
    t.fn[PREC][1][1][1] = &launch_backward<PREC, 1, 1, 128>;
    t.strided[PREC][0][0][0] = &launch_strided<PREC, 0, 0, 64>;
    t.strided[PREC][0][1][0] = &launch_strided<PREC, 0, 1, 64>;
    t.strided[PREC][1][0][0] = &launch_strided<PREC, 1, 0, 64>;
    t.strided[PREC][1][1][0] = &launch_strided<PREC, 1, 1, 64>;
    t.strided[PREC][0][0][1] = &launch_strided<PREC, 0, 0, 128>;
    t.strided[PREC][0][1][1] = &launch_strided<PREC, 0, 1, 128>;
    t.strided[PREC][1][0][1] = &launch_strided<PREC, 1, 0, 128>;
    t.strided[PREC][1][1][1] = &launch_strided<PREC, 1, 1, 128>;
    t.conv_fwd[PREC][0] = &launch_conv_fwd<PREC, 64>;
    t.conv_fwd[PREC][1] = &launch_conv_fwd<PREC, 128>;
    t.conv_bwd[PREC][0][0] = &launch_conv_backward<PREC, 64, false>;
    t.conv_bwd[PREC][1][0] = &launch_conv_backward<PREC, 128, false>;
    t.conv_bwd[PREC][0][1] = &launch_conv_backward<PREC, 64, true>;
    t.conv_bwd[PREC][1][1] = &launch_conv_backward<PREC, 128, true>;
}

void register_backward_fp32(BwdTable &t);
void register_backward_bf16x3(BwdTable &t);
void register_backward_bf16(BwdTable &t);
void register_backward_bf16x6(BwdTable &t);

}  // namespace ls
