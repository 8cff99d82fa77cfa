// Shared definitions of the MI355X (gfx950) matrix-core FFT library: data formats,
// precision modes and the argument block of the tile kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfft {

// Complex element formats at a kernel boundary (interleaved re, im).
enum Fmt : int {
    FMT_C16 = 0,   // 2 x fp16  (4 B)  -- the reference's half2 layout (tcfft_half.h:20)
    FMT_CBF16 = 1, // 2 x bf16  (4 B)
    FMT_C32 = 2,   // 2 x fp32  (8 B)
};

// Arithmetic of the radix-16 stages.
enum Prec : int {
    PREC_FP16 = 0,  // fp16 MFMA operands, fp32 accumulate
    PREC_BF16 = 1,  // bf16 MFMA operands, fp32 accumulate
    PREC_SPLIT = 2, // split fp16 hi+lo operands (3 products), fp32 accumulate
    PREC_FP32 = 3,  // plain fp32 VALU butterflies (the comparator)
};

constexpr int kMaxStages = 8;
constexpr int kWaves = 4;                 // waves per workgroup of the tile kernel
constexpr int kThreads = kWaves * 64;
constexpr int kMaxTile = 4096;            // complex points per workgroup tile
constexpr int kMaxPairsPerWave = kMaxTile / 256 / kWaves;  // MFMA column groups per wave/stage
constexpr int kMaxValuPerThread = 32;
// Cache policy of the 2048-point on-chip line kernels' output stores: nt (non-temporal, bit
// 1 of the buffer instructions' policy operand on gfx950). Their outputs are a plan's final
// result, which no later pass of the plan reads, and each store instruction writes whole
// 128-byte row segments: 2048^2 x 16 columns 71.7 -> 67.2 us per launch, 2^20 x 64 rows 69.3
// -> 65.1 us, HBM traffic unchanged. Kept off elsewhere: the 65536-point kernel's 64-byte store
// segments then wrote 14% more bytes for no measurable gain, multi-pass kernels' intermediates
// are read by the next pass from the caches (with nt, 1024^2 x 64 2D ran 9% and the 2048^2
// conv 5% slower), and nt loads made the 65536-point kernel 7% slower.
constexpr int kStoreStream = 2;
constexpr int kFastTw = 256;              // four-step table entries staged in LDS (each half)
constexpr int kFastTwCols = 1024;         // the same for the strided kind (the four-step's first
                                          // pass, where the twiddles are applied; N <= 2^20)

__host__ __device__ constexpr int lds_pad(int e) { return e + (e >> 4); }
__host__ __device__ constexpr int tile_lds_bytes(int points) { return lds_pad(points) * 8 + 64; }

// Mixed-radix digit reversal of an index whose radices are powers of two: digit i of x
// (least significant first) is bits [sh[i], sh[i] + log2 radix) and moves to bit position
// out[i]. n = 0 is the identity.
struct DigitRev {
    int n;
    int sh[4], mask[4], out[4];
};

__host__ __device__ inline int digit_rev(int x, const DigitRev &d)
{
    if (d.n == 0) return x;
    int r = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < d.n) r |= ((x >> d.sh[i]) & d.mask[i]) << d.out[i];
    return r;
}

// One workgroup transforms T lines of R points. Lines are indexed by (l0, l1, l2);
// l0 runs fastest and a tile covers T consecutive l0. Global element (line, point)
// lives at  l0*s0 + l1*s1 + l2*s2 + point*ps  (in complex elements) on each side.
struct TileArgs {
    const void *in;
    void *out;
    int in_fmt, out_fmt;
    int R;                     // points per line
    int nstages;
    int radix[kMaxStages];     // stage radices, product = R (16 -> matrix-core stage)
    int T;                     // lines per tile (divides c0)
    int TL;                    // lines interleaved in one MFMA column group (1..16)
    int c0, c1, c2;            // line counts per level
    long long in_s0, in_s1, in_s2, in_ps;
    long long out_s0, out_s1, out_s2, out_ps;
    const int *in_perm;        // optional: global point i holds logical point in_perm[i]
    const float2 *stage_tw;    // W_R^e, e in [0, R), sign of the transform
    const void *dft_ops;       // per-lane DFT16 operands (see DftOps)
    // four-step twiddle after the last stage: out *= W_Ntw^(l0 * k), e = l0*k < Ntw
    int tw_N;                  // 0 = no pass twiddle
    int tw_level;              // 0: exponent l0 * k ; 1: exponent l1 * k
    int tw_bits;               // W_N^e = tw_hi[e >> bits] * tw_lo[e & ((1<<bits)-1)]
    const float2 *tw_lo, *tw_hi;
    float stage_mul[kMaxStages];  // multiplier folded into each stage (1 or 1/radix)
    float out_mul;             // extra multiplier applied at the store
    int sign;                  // -1 forward, +1 inverse (VALU stages)
    int in_vec, out_vec;       // base pointers are 16-byte aligned (vector copies allowed)
    // fast kernels: scale folded into the tables
    const float2 *stage_tw_s;  // W_R^e * (per-stage scale, / 16 for matrix-core operands)
    const float2 *tw_lo_s;     // tw_lo * last-stage scale
    float last_mul;            // last-stage scale when there is no pass twiddle
    float tail_mul;            // scale of the radix-2/4/8 stage
    int tw_lds;                // four-step tables fit in LDS (2^tw_bits and the hi part <= kFastTw,
                               // kFastTwCols for the strided kind)
    // fused first-pass input (streaming kernel only; zero/null otherwise), element e of the
    // chunk = transform batch_base + e / real_p, offset e % real_p inside it:
    //   in_mul: the loaded value is multiplied by in_mul[e % in_mul_period]
    //   real_src: the value is (img[2k][c][y][x], img[2k+1][c][y][x]) of real images
    //   [real_b][real_c][real_h][real_w] (k = transform / real_c, c = transform % real_c;
    //   zero outside the image: the padding of the FFT convolution)
    const float2 *in_mul;
    long long in_mul_period;
    const void *real_src;
    const void *real_src2;     // set: the imaginary part comes from this second image instead
    int real_f64, real_b, real_c, real_h, real_w, real_pw;
    long long real_p, batch_base;
    // fast kernels, digit-reversed input (the order of tcfftExec, gen_rev of a radix list
    // whose first radix is 16): natural point col + (R/16) j of a line is stored at point
    // position j + 16 * digit_rev(col, in_rev), so each lane's four first-stage operands
    // are consecutive points. in_rev.n = 0 with rev_input = 0: natural order.
    int rev_input;
    DigitRev in_rev;
    // four-step twiddle of a digit-reversed first pass: the exponent's line factor is
    // digit_rev(line, tw_rev) (identity when tw_rev.n = 0)
    DigitRev tw_rev;
    // cos(pi/8), sin(pi/8), cos(pi/4): the constants of the VALU radix-8 and radix-16
    // butterflies, generated in FP64 on the GPU and rounded once to fp32 like every other
    // table (gen_w16)
    float w16[3];
    // streaming kernel: bit 0 (1) every input byte offset of the launch, bit 1 (2) every
    // output byte offset, is below 2^32; in_span / out_span: the bytes the launch reaches from
    // in / out (the extent of the buffer resource its loads / stores then go through)
    int off32;
    unsigned in_span, out_span;
};

// Threads of a streaming launch with the fused filter product (TileArgs::in_mul): the lines
// of a transform in whole groups of 64, times the launch's transforms.
__host__ __device__ inline long long in_mul_threads(long long lines, long long transforms)
{
    return (lines + 63) / 64 * 64 * transforms;
}

// DFT16 operands for the A side of v_mfma_f32_16x16x32_{f16,bf16}: per lane 8 values of
// [Fr | -s*Fi] (real output rows) and [s*Fi | Fr] (imag output rows), scaled by 16.
struct alignas(16) DftOps {
    _Float16 hr[64][8], hi_[64][8], lr[64][8], li[64][8];  // fp16 hi/lo, real and imag rows
    __bf16 br[64][8], bi[64][8];                          // bf16 real and imag rows
};

// ------------------------------------------------------------------ fp16 / bf16 packing
typedef float pk_f2 __attribute__((ext_vector_type(2)));
typedef _Float16 pk_h2 __attribute__((ext_vector_type(2)));
typedef __bf16 pk_b2 __attribute__((ext_vector_type(2)));

// (re, im) -> one word of two fp16 / bf16 in one packed conversion (round to nearest even,
// the same result as two scalar conversions).
__device__ __forceinline__ uint32_t pack_c16(float re, float im)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((pk_f2){re, im}, pk_h2));
}
__device__ __forceinline__ uint32_t pack_cbf16(float re, float im)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((pk_f2){re, im}, pk_b2));
}

// Split fp32 (re, im) into fp16 pairs hi = fp16(x), lo = fp16(x - hi); x - hi is exact in
// fp32. The difference is formed as fma(hi, m, x) with m = -1 hidden from the optimiser by
// an empty asm statement, so it compiles to v_fma_mix_f32 reading the fp16 half in place
// instead of a conversion back to fp32 and a subtraction. (Instructions written in inline
// asm would get no hazard wait states from the compiler: one that reads an MFMA result right
// after the MFMA reads a stale accumulator.)
__device__ __forceinline__ void split_c16(float re, float im, pk_h2 &hi, pk_h2 &lo)
{
    hi = __builtin_convertvector((pk_f2){re, im}, pk_h2);
    float m = -1.0f;
    __asm__("" : "+s"(m));
    lo = __builtin_convertvector((pk_f2){__builtin_fmaf((float)hi.x, m, re), __builtin_fmaf((float)hi.y, m, im)},
                                 pk_h2);
}

}  // namespace mfft
