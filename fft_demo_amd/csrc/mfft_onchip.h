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

// The kernel for a precision and its 16-bit format, or null. rev: the lines are stored in the
// digit-reversed order of four radix-16 digits (split fp16 and fp16 only).
const void *find_onchip_kernel(int prec, int fmt, bool rev);

}  // namespace mfft
