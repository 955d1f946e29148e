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
#include <cstddef>
#include <cstdint>

#include "gemm_core.hpp"

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

// MFMA cycles one K chunk of a tile costs on its busiest SIMD (waves w and
// w + 4 of an eight-wave tile share SIMD w % 4), for a GEMM covering mt x nt
// 16 x 16 blocks of it.  FP32 / FP64 waves multiply the sub-tiles that reach
// into the GEMM; the split modes' tile kernels every sub-tile of a wave
// holding part of it.  The wave placement is the kernels' own
// (TileCfg::wave_g, rows_in, ...).
template <class C>
inline double busiest_simd(long M, long N)
{
    const int mt = (int)cdiv(std::min<long>(C::BM, M), 16), nt = (int)cdiv(std::min<long>(C::BN, N), 16);
    double simd[4] = {0, 0, 0, 0};
    for (int w = 0; w < C::NW; ++w) {
        const long rows = C::rows_in(C::wave_g(w), mt), cols = C::cols_in(C::wave_h(w), nt);
        simd[w % 4] += rows * cols == 0 ? 0.0 : C::INTERLEAVE ? (double)rows * cols : (double)C::WM * C::WN;
    }
    return *std::max_element(simd, simd + 4);
}

// The same for any tile shape; the placement depends only on whether the
// mode splits its inputs (PREC stands for its class).
template <int PREC>
inline double busiest_simd_tile(int tile, long M, long N)
{
    switch (tile) {
    case T32x32: return busiest_simd<TileCfg<PREC, 0, 0, 32, 32>>(M, N);
    case T64x64: return busiest_simd<TileCfg<PREC, 0, 0, 64, 64>>(M, N);
    case T128x128: return busiest_simd<TileCfg<PREC, 0, 0, 128, 128>>(M, N);
    case T128x32: return busiest_simd<TileCfg<PREC, 0, 0, 128, 32>>(M, N);
    case T128x64: return busiest_simd<TileCfg<PREC, 0, 0, 128, 64>>(M, N);
    default: return busiest_simd<TileCfg<PREC, 0, 0, 32, 128>>(M, N);
    }
}

inline double tile_chunk_mfma(int prec, int tile, long M, long N)
{
    const bool split = prec == P_BF16X3 || prec == P_BF16 || prec == P_BF16X6;
    const double busiest = split ? busiest_simd_tile<P_BF16X3>(tile, M, N) : busiest_simd_tile<P_FP32>(tile, M, N);
    return busiest * subtile_chunk_cycles(prec);
}

// Estimated cycles for one GEMM with a given tile and split: rounds of
// resident workgroups, each a pipeline of K chunks bound by the busier of
// its MFMA work and its share of the CU's memory bandwidth (co-resident
// workgroups share both), plus, with K slices, the last slice's sum of the
// partial tiles.
inline double plan_cost(int prec, long M, long N, long K, int tile, int ksplit, int num_cu)
{
    const TileShape ts = kTileShapes[tile];
    const long tm = cdiv(M, ts.bm), tn = cdiv(N, ts.bn);
    const long wgs = tm * tn * ksplit;
    const int occ = tile_occupancy(prec, tile);
    const double rounds = (double)cdiv(wgs, (long)num_cu * occ);
    const double co = (double)std::min<long>(occ, cdiv(wgs, num_cu));  // workgroups sharing a CU
    const double rows = std::min<double>(ts.bm, (double)M), cols = std::min<double>(ts.bn, (double)N);
    const int esize = prec == P_FP64 ? 8 : 4;
    const double mfma = tile_chunk_mfma(prec, tile, M, N) * co;
    // The constants (cycles per workgroup and per chunk, bytes per cycle a
    // CU's loads sustain, the last slice's cost per partial-tile byte) were
    // fitted to measured times of every tile and K split on the tall-skinny
    // and short-wide sgemm_mp grid (M or N in {4096, 16384, 65536}, the other
    // in {16, 32, 64}, K in {64, 256, 1024}; FP32 and BF16X3): the plans they
    // choose take 3% longer than the fastest measured plan of each shape.
    const double mem = (rows + cols) * chunk_depth(prec) * esize * co / 40.0;
    const long kslice = cdiv(K, ksplit);
    const double chunks = (double)std::max<long>(1, cdiv(kslice, chunk_depth(prec)));
    double total = rounds * (12000.0 + chunks * (std::max(mfma, mem) + 150.0));
    if (ksplit > 1)  // partial tile written, then read back by the last slice
        total += 2500.0 + rows * cols * esize * (1.0 + ksplit) / 40.0;
    return total;
}

// The plan of one GEMM: the tile shape and power-of-two K split of least
// estimated cost, with at most max_slot_bytes of partial tiles and
// max_split_tiles tiles when K is split (0: no split).
inline SinglePlan plan_single(int prec, long M, long N, long K, bool /*mp*/, int num_cu,
                              size_t max_slot_bytes = (size_t)256 << 20, size_t max_split_tiles = 1 << 17)
{
    SinglePlan best{T32x32, 1, (int)std::max<long>(K, 1)};
    double best_cost = 1e300;
    const int kd = chunk_depth(prec);
    const int esize = prec == P_FP64 ? 8 : 4;
    for (int t = 0; t < NTILE; ++t) {
        if (prec == P_FP64 && t == T128x128) continue;  // 128 accumulator registers: occupancy 1
        // FP32: no 128 x 128 tile for a GEMM of at most 64 rows or columns.
        // Half of its eight waves would hold nothing, yet run the chunk
        // pipeline (loads, LDS writes, barriers) with the others.  The MFMA
        // model alone ties it with the four-wave tall / wide tiles there
        // (the busiest SIMD carries the same sub-tiles), and measured it is
        // slower (4096 x 16 x 1024: 15.7 us against 10-12 us).
        if (prec == P_FP32 && t == T128x128 && std::min(M, N) <= 64) continue;
        const TileShape ts = kTileShapes[t];
        const size_t ntiles = (size_t)cdiv(M, ts.bm) * (size_t)cdiv(N, ts.bn);
        for (int ks = 1; ks <= 64; ks *= 2) {
            if (ks > 1 && K < (long)ks * 4 * kd) break;  // keep >= 4 chunks per slice
            if (ks > 1 && (ntiles > max_split_tiles || ntiles * ks * ts.bm * ts.bn * esize > max_slot_bytes)) break;
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
