// On-chip kernel for 65536-point lines (one workgroup per line; mfft_onchip.hip).
#pragma once
#include "mfft_common.h"

namespace mfft {

// Lines of 65536 points, natural order, 16-bit complex words on both sides (the format of
// the precision: fp16 pairs for split fp16 and fp16, bf16 pairs for bf16).
struct OnchipArgs {
    const void *in;
    void *out;
    long long batch;               // lines
    long long in_dist, out_dist;   // words between consecutive lines
    const void *dft_ops;           // DftOps of the transform's direction
    const float2 *stage_tw_s;      // W_256^e * fac, e in [0, 256)
    const float2 *tw_lo_s;         // W_65536^e * fac, e in [0, 256)
    const float2 *tw_hi;           // W_65536^(256 e), e in [0, 256)
    float fac;                     // matrix-core stage scale (per-stage 1/16 or 1, / 16)
};

constexpr int kOnchipN = 65536;
constexpr int kOnchipThreads = 512;

// 2048-point lines, 32 per workgroup, transformed whole, 16-bit words on both sides:
//   columns: the column axis of a 2D transform with 2048 rows, in place in the same layout
//            (column c, point n at c + n * ps on both sides);
//   rows:    32 consecutive contiguous lines (line l, point n at 2048 l + n), stored
//            transposed (point k of line l at l + k * ps_out): the second pass of a 1D
//            four-step plan N = L1 x 2048 (ps_out = L1).
struct OnchipColsArgs {
    const void *in;
    void *out;
    long long blocks;              // workgroups: 32-line blocks per transform x transforms
    int col_blocks;                // 32-line blocks per transform
    int ps;                        // columns: words between consecutive points (the row pitch)
    int ps_out;                    // rows: words between consecutive output points (L1)
    long long in_dist, out_dist;   // words between consecutive transforms
    const void *dft_ops;           // DftOps of the transform's direction
    const float2 *tw_s;            // W_2048^e * fac, e in [0, 2048)
    const float2 *stw_s;           // W_128^e * fac, e in [0, 128)
    float tail_mul;                // radix-8 stage scale (1/8 or 1)
    int sign;
    float w16[3];                  // cos(pi/8), sin(pi/8), cos(pi/4)
};

constexpr int kOnchipColsN = 2048;
constexpr int kOnchipColsLines = 32;

const void *find_onchip_cols_kernel(int prec, int fmt, bool rows = false);

// The kernel for a precision and its 16-bit format, or null. rev: the lines are stored in the
// digit-reversed order of four radix-16 digits (split fp16 and fp16 only).
const void *find_onchip_kernel(int prec, int fmt, bool rev);

}  // namespace mfft
