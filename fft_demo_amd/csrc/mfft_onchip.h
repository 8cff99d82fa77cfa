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

// Strided 2048-point lines (the column axis of a 2D transform with 2048 rows): one
// workgroup transforms 32 adjacent columns whole, in place in the same layout (column c,
// point n at c + n * ps). 16-bit words on both sides.
struct OnchipColsArgs {
    const void *in;
    void *out;
    long long blocks;              // workgroups: (columns / 32) per image x images
    int col_blocks;                // 32-column blocks per image
    int ps;                        // words between consecutive points of a column (the row pitch)
    long long in_dist, out_dist;   // words between consecutive images
    const void *dft_ops;           // DftOps of the transform's direction
    const float2 *tw_s;            // W_2048^e * fac, e in [0, 2048)
    const float2 *stw_s;           // W_128^e * fac, e in [0, 128)
    float tail_mul;                // radix-8 stage scale (1/8 or 1)
    int sign;
    float w16[3];                  // cos(pi/8), sin(pi/8), cos(pi/4)
};

constexpr int kOnchipColsN = 2048;
constexpr int kOnchipColsLines = 32;

const void *find_onchip_cols_kernel(int prec, int fmt);

// The kernel for a precision and its 16-bit format, or null. rev: the lines are stored in the
// digit-reversed order of four radix-16 digits (split fp16 and fp16 only).
const void *find_onchip_kernel(int prec, int fmt, bool rev);

}  // namespace mfft
