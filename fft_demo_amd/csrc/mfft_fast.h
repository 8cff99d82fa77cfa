// Fast tile kernels of the matrix-core FFT (power-of-two line lengths 64..4096; the
// strided kind 64..1024).
//
// Same method as the generic tile kernel (radix-16 DFT-matrix products on MFMA, Stockham
// schedule, LDS between stages) with the HBM traffic moved into the stages themselves:
// the first stage loads its MFMA operands straight from global memory, the last stage
// stores its accumulators straight to global memory (with the four-step twiddle and the
// output conversion fused), and LDS is touched only between stages. Geometry, formats
// and stage count are compile-time, so index arithmetic folds to shifts.
//
// Kinds (how a tile of T lines x R points maps onto memory):
//   KIND_ROWS     points contiguous (in_ps = out_ps = 1); lanes walk the points of a line.
//   KIND_ROWS_TR  points contiguous on input, lines contiguous on output (out_s0 = 1):
//                 the last stage switches the lanes to walk the lines.
//   KIND_COLS     lines contiguous on both sides (in_s0 = out_s0 = 1, T = 8192 / R lines);
//                 lanes walk the lines of the tile
//                 (strided passes: the first pass of the four-step method, the columns
//                 of a 2D transform).
#pragma once
#include "mfft_common.h"

namespace mfft {

enum { KIND_ROWS = 0, KIND_ROWS_TR = 1, KIND_COLS = 2, KIND_ONCHIP = 3, KIND_ONCHIP_COLS = 4 /* mfft_onchip.h */ };
// Lines per tile of the strided kind: R * T = 8192 points at every length (128 lines at 64
// points, 64 at 128, 32 at 256, 16 at 512, 8 at 1024). A 16-wave workgroup then keeps two
// 16-column MFMA groups per wave, whose accumulators, stage values and next-tile operands fit
// the 128 registers a wave has at 4 waves per SIMD (four groups per wave spill to scratch).
// Each point of a tile is one segment of T consecutive lines in memory, and a strided copy on
// MI355X streams 1.3-1.6 TB/s with 32-byte segments, 3.0-3.1 with 64, 4.5-4.9 with 128 and
// 5.0-5.2 with 256 bytes (5.7 for whole rows): short lines get more of them per tile.
constexpr int kColsMaxR = 1024;
constexpr int cols_lines(int R) { return 8192 / R; }

// Waves per workgroup (host and device). fp32: one 16-point column per thread in the radix-16
// stages. Matrix-core precisions: the strided kind keeps two 16-column MFMA groups per wave
// (16 waves for its 8192-point tiles); the row kinds keep one group per wave (16 waves) for
// lines of 512 points or more and two (8 waves) below. Measured on 2^26-point batches, one
// group per wave ran the row kernels 1-5% faster at 512..4096 points (the 1024^2 x 64 2D
// exec 265 -> 254 us); at 256 points and 4096 lines (one tile per compute unit) it ran 24%
// slower (4.3 -> 5.3 us per exec).
constexpr int fast_waves_of(int prec, int R, int kind)
{
    const int T = kind == KIND_COLS ? cols_lines(R) : kMaxTile / R;
    const int groups = R * T / 256;
    if (prec == PREC_FP32) return R * T / 16 / 64;
    const int per_wave = (kind != KIND_COLS && R >= 512) ? 1 : 2;
    const int w = groups / per_wave;
    return w < 4 ? 4 : (w > 16 ? 16 : w);
}

// Precisions and input formats whose fast kernels also read digit-reversed lines
// (TileArgs::rev_input): the matrix-core paths with fp16 or fp32 input.
constexpr bool fast_rev_ok(int prec, int in_fmt) { return prec != PREC_FP32 && in_fmt != FMT_CBF16; }

typedef void (*fast_fn)(TileArgs);
fast_fn find_fast_kernel(int prec, int R, int kind, int in_fmt, int out_fmt);
int fast_tile_lines(int R, int kind);
int fast_threads(int prec, int R, int kind);

}  // namespace mfft
