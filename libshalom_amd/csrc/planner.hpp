// planner.hpp -- host-side tile / split-K planner.
//
// Replaces the reference's OpenMP partition planner (Dete_grad_N_threads_nums /
// Dete_grad_M_threads_nums, SMM_thread.c:14-72, which shapes a Tm x Tn thread
// grid after the M/N aspect ratio) with a choice of workgroup tile shape and
// K-split for the GPU: a small analytic model of one tile's time (MFMA cycles
// vs. its share of HBM bandwidth per K chunk) times the number of rounds of
// workgroups the chip runs, minimised over the tile shapes and power-of-two
// K splits.  Host code only: it is compiled into the library and exercised by
// the CPU tests through LibShalom_query_plan.
#pragma once
#include <algorithm>
#include <cstdint>

namespace ls {

struct SinglePlan {
    int tile;    // TileId
    int ksplit;  // number of K slices (1 = no split)
    int kper;    // K extent of one slice (multiple of the chunk depth)
};

// Workgroups per CU of the single-GEMM tile kernel for each (precision,
// tile), from its LDS and register use on gfx950 (160 KB LDS, 512 VGPR+AGPR
// per SIMD lane; four waves per workgroup, eight for 128 x 128).
inline int tile_occupancy(int prec, int tile)
{
    //                                 32x32 64x64 128x128 128x32 128x64 32x128
    static const int fp32[NTILE] = {7, 4, 1, 3, 2, 3};
    static const int bf16[NTILE] = {8, 4, 2, 4, 2, 4};
    static const int bf3[NTILE] = {8, 4, 1, 4, 2, 4};
    static const int bf6[NTILE] = {6, 3, 1, 2, 2, 2};
    static const int fp64[NTILE] = {7, 3, 1, 3, 2, 3};
    switch (prec) {
    case P_FP32: return fp32[tile];
    case P_BF16: return bf16[tile];
    case P_BF16X3: return bf3[tile];
    case P_BF16X6: return bf6[tile];
    default: return fp64[tile];
    }
}

inline int chunk_depth(int prec) { return prec == P_FP64 ? 16 : 32; }

// MFMA cycles for one 16x16 output sub-tile over one K chunk.
inline double subtile_chunk_cycles(int prec)
{
    switch (prec) {
    case P_FP32: return 8 * 32.0;   // 8 x v_mfma_f32_16x16x4_f32
    case P_BF16: return 16.0;       // 1 x v_mfma_f32_16x16x32_bf16
    case P_BF16X3: return 3 * 16.0;
    case P_BF16X6: return 8 * 16.0;
    default: return 4 * 64.0;       // 4 x v_mfma_f64_16x16x4_f64
    }
}

inline long cdiv(long a, long b) { return (a + b - 1) / b; }

// Estimated cycles for one GEMM with a given tile and split.
inline double plan_cost(int prec, long M, long N, long K, int tile, int ksplit, int num_cu)
{
    const TileShape ts = kTileShapes[tile];
    const long tm = cdiv(M, ts.bm), tn = cdiv(N, ts.bn);
    const long wgs = tm * tn * ksplit;
    const int occ = tile_occupancy(prec, tile);
    const long slots = (long)num_cu * occ;
    const double rounds = (double)cdiv(wgs, slots);
    // Average rows / cols of C a tile really covers (16-granular).
    const double rows = std::min<double>(ts.bm, cdiv(M, 16) * 16.0);
    const double cols = std::min<double>(ts.bn, cdiv(N, 16) * 16.0);
    const double subtiles_per_wave = std::max(1.0, (rows / 16.0) * (cols / 16.0) / 4.0);
    const double mfma = subtiles_per_wave * subtile_chunk_cycles(prec) * std::min<double>(occ, 4.0);
    const int esize = prec == P_FP64 ? 8 : 4;
    const double bytes = (rows + cols) * chunk_depth(prec) * esize;
    // ~8 TB/s over the CUs at 2.4 GHz ~ 13 bytes per cycle per CU, shared by
    // the co-resident workgroups.
    const double mem = bytes * std::min<double>((double)occ, (double)cdiv(wgs, num_cu)) / 13.0;
    const long kslice = cdiv(K, ksplit);
    const double chunks = (double)std::max<long>(1, cdiv(kslice, chunk_depth(prec)));
    const double per_wg = 1500.0 + chunks * (std::max(mfma, mem) + 150.0);
    double total = rounds * per_wg;
    if (ksplit > 1) total += 3000.0 + (double)ksplit * M * N * esize / (13.0 * num_cu);
    return total;
}

inline SinglePlan plan_single(int prec, long M, long N, long K, bool /*mp*/, int num_cu)
{
    SinglePlan best{T32x32, 1, (int)std::max<long>(K, 1)};
    double best_cost = 1e300;
    const int kd = chunk_depth(prec);
    for (int t = 0; t < NTILE; ++t) {
        if (prec == P_FP64 && t == T128x128) continue;  // 128 accumulator registers: occupancy 1
        // The 128 x 128 tile's eight waves own fixed 32 x 64 blocks: a GEMM
        // with a dimension of at most 64 leaves half of them idle, which the
        // FP32 MFMAs (the slowest per element) cannot afford; the tall and
        // wide tiles split the rows (columns) over all their waves instead.
        if (prec == P_FP32 && t == T128x128 && std::min(M, N) <= 64) continue;
        for (int ks = 1; ks <= 64; ks *= 2) {
            if (ks > 1 && K < (long)ks * 4 * kd) break;  // keep >= 4 chunks per slice
            const double c = plan_cost(prec, M, N, K, t, ks, num_cu);
            if (c < best_cost * 0.999) {
                best_cost = c;
                const long kper = ks == 1 ? std::max<long>(K, 1) : cdiv(cdiv(K, ks), kd) * kd;
                best = SinglePlan{t, ks == 1 ? 1 : (int)cdiv(K, kper), (int)kper};
            }
        }
    }
    return best;
}

// Tile for one GEMM of a batch: the smallest square tile that covers it
// (whole GEMM in one workgroup when it is at most 128 x 128).
inline int batched_tile(int prec, long M, long N, long /*K*/)
{
    const long mx = std::max(M, N);
    if (mx <= 32) return T32x32;
    if (mx <= 64 || prec == P_FP64) return T64x64;
    return T128x128;
}

}  // namespace ls
