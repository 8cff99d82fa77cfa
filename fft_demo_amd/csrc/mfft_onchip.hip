// On-chip kernel for 65536-point lines: one workgroup transforms a whole line, keeping the
// 256 x 256 intermediate of the four-step method on the chip, so each line is read from and
// written to HBM once (the two-pass plan also writes and reads the intermediate: twice the
// traffic).
//
// The method is the reference's four-step decomposition of a long line
// (tcfft_half.cu:266-390, layer_256_1 followed by layer_256_0), on the MFMA DFT16 of the
// fast kernels:
//   x[256 n1 + n2]: phase 1, for each column n2: DFT256 over n1 (two radix-16 stages),
//                   times W_65536^(n2 k1) -> Y[k1][n2];
//                   phase 2, for each row k1: DFT256 over n2 -> X[k1 + 256 k2].
// Storage: Y is 65536 16-bit complex words (256 KB) but a compute unit has 160 KB of LDS,
// so phase 2 runs in four rounds of 64 rows. Phase 1 writes the rows of round 0 to LDS and
// keeps the other three rounds in registers (96 words per lane of 512 lanes); round r
// (r >= 1) first dumps its words into the LDS rows. The DFT16 operand rows are permuted so a
// lane's four outputs of a group fall one in each round (output row 4g + i holds frequency
// 4i + g), which spreads the held words evenly over the lanes.
//
// Geometry: 8 waves (2 per SIMD, 256 registers each). A stage transforms 32 lines (columns
// in phase 1, rows in phase 2) = 512 DFT16s = 32 MFMA groups, 4 per wave. Between the two
// radix-16 stages of a DFT256 the values go through an fp32 exchange buffer (64 KB). Phase-1
// loads run one 32-column chunk ahead of the compute (16 words per lane). Global loads and
// stores go through a buffer resource per line (scalar base, 32-bit lane offsets).
#include "mfft_fast_impl.h"
#include "mfft_onchip.h"

namespace mfft {
namespace onchip {
using fast::f4;
using fast::h8;
using fast::b8;
using fast::Ops;
using fast::dft16;
using fast::cmulf;
using fast::Elem;

constexpr int NW = 8;
constexpr int NT = NW * 64;
constexpr int SIDE = 256;        // x[256 n1 + n2]
constexpr int CH = 32;           // lines per stage (phase-1 columns / phase-2 rows)
constexpr int NCH = SIDE / CH;   // 8 column chunks
constexpr int GPW = 4;           // MFMA groups per wave per stage
constexpr int ROWS = 64;         // rows of Y in LDS at a time (one round)
constexpr int ROUNDS = SIDE / ROWS;
// Phase-1 chunks in flight ahead of the compute. One (32 KB per compute unit) measured
// 1.3 us faster on 65536 x 256 than two or three: the loads are not what limits the kernel
// (without them it runs only 3 us faster), and the deeper buffers cost registers.
#ifndef MFFT_ONCHIP_DEPTH
#define MFFT_ONCHIP_DEPTH 1
#endif
constexpr int DEPTH = MFFT_ONCHIP_DEPTH;

// Y rows in LDS: word (row, n2) at row * YP + n2 + 16 (n2 >> 6). With YP = 1 (mod 64) the
// lanes of a write (16 consecutive n2 x 4 rows 16 apart) and of a read (16 consecutive rows
// x 4 points 64 apart) fall on 64 distinct banks, and every address is a per-lane base plus
// a compile-time offset (no per-access index arithmetic, so no address registers held).
constexpr int YP = SIDE + 3 * 16 + 17;  // 321 words per row
static_assert(YP % 64 == 1 && YP >= SIDE + 3 * 16, "row pitch");

// Exchange buffer between the radix-16 stages: (a, f, t) = (input digit of the second
// stage, frequency of the first, line), float2 at [16 a + f][t ^ 16 b] with b = bit 0 of
// (row ^ row >> 6). Writes (16 consecutive t x 4 consecutive f) and reads (16 consecutive t
// x 4 values of a, 4 apart) then hit 32 distinct float2 slots per half wave. For the fixed
// lane and group patterns the swizzle bit is constant per lane (writes: (g ^ u) & 1, reads:
// (g ^ (wave >> 1)) & 1), so both are per-lane bases plus offsets.
__host__ __device__ constexpr int ephys(int a, int f, int t)
{
    return (16 * a + f) * CH + (t ^ ((((16 * a + f) ^ ((16 * a + f) >> 6)) & 1) << 4));
}

// W_65536^e from the staged tables: lo (scaled by the stage factor) at [0, 256), hi at
// [256, 512).
__device__ __forceinline__ float2 tw65536(const float2 *twl, unsigned e)
{
    e &= 65535u;
    return cmulf(twl[e & 255u], twl[256 + (e >> 8)]);
}

template <int PREC, int FMT>
struct Chunk {
    typedef typename Elem<FMT>::raw raw;
    raw w[GPW * 4];
};

// Stage 1 of a DFT256 over 32 lines: group u of this wave transforms digit a = (wave >> 1) +
// 4u of lines t = 16 (wave & 1) + c; the lane's four inputs are points a + 16 (4g + i) in the
// 16-bit words `w` (exact MFMA operands). Results times W_256^(a f) (scaled) go to the
// exchange buffer at (a, f, t), f = 4i + g. The caller has passed a barrier since the
// previous stage 2 read the buffer; each group's products are written as they complete, so
// only one group's accumulators are live.
template <int PREC, int FMT>
__device__ __forceinline__ void stage1(const typename Elem<FMT>::raw *w, const Ops<PREC> &op, const float2 *stw,
                                       float2 *ex, int w2, int g, const int *exw)
{
#pragma unroll
    for (int u = 0; u < GPW; ++u) {
        fast::u4 v = {w[4 * u], w[4 * u + 1], w[4 * u + 2], w[4 * u + 3]};
        const float2 x[4] = {};
        f4 ar, ai;
        dft16<PREC, true>(x, __builtin_bit_cast(h8, v), __builtin_bit_cast(b8, v), op, ar, ai);
        const int a = w2 + 4 * u;
        const float2 *sw = stw + a * g;  // W_256^(a f), f = 4i + g: index a g + 4 a i
#pragma unroll
        for (int i = 0; i < 4; ++i)
            ex[exw[u & 1] + (64 * u + 4 * i) * CH] = cmulf(make_float2(ar[i], ai[i]), sw[4 * a * i]);
    }
    __syncthreads();
}

// Stage 2, group u: transforms frequency f = (wave >> 1) + 4u of lines t = 16 (wave & 1) + c
// over a; the lane reads a = 4g + i and gets outputs v = 4i + g (point f + 16 v of the line).
// exr: the lane's read base (a = 4g, f = wave >> 1).
template <int PREC>
__device__ __forceinline__ void stage2(const float2 *ex, const Ops<PREC> &op, int u, int exr, f4 &ar, f4 &ai)
{
    float2 x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = ex[exr + (16 * i + 4 * u) * CH];
    dft16<PREC, false>(x, h8{}, b8{}, op, ar, ai);
}

// Buffer resource of one line (gfx9 descriptor word 3; the record count is the line's bytes,
// so the hardware drops any access outside the line). Loads and stores through it take the
// lane's 32-bit byte offset and a uniform byte offset in a scalar register: no 64-bit address
// arithmetic per access.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t line_rsrc(const void *base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, kOnchipN * 4, 0x00020000);
}

// Chunk j of phase 1: point n1 = a + 16 (4g + i) of column n2 = 32 j + t. 16 lanes read 16
// consecutive columns (64 bytes), the two waves of a pair the two halves of a 128-byte row
// segment; `wave` is uniform, so the row part of the offset is scalar.
//
// REV: the line is stored in the digit-reversed order of four radix-16 digits (gen_rev of
// {16, 16, 16, 16}, tcfft_half.cu:1351-1374: the order tcfftExec reads), where natural point
// 256 n1 + n2 sits at rev(n1) + 256 rev(n2), rev swapping the two hex digits of a byte. The
// lane's four points n1 = a + 16 (4g + i) then sit at 4g + i + 16 a + 256 rev(n2): one
// 16-byte load, with rev(32 j + 16 (wave & 1) + c) = 2 j + (wave & 1) + 16 c. A chunk reads
// 32 whole 1 KB rows of the stored line.
template <int PREC, int FMT, bool REV>
__device__ __forceinline__ void load_chunk(Chunk<PREC, FMT> &ck, __amdgpu_buffer_rsrc_t src, int j, int wave,
                                           unsigned lane_off)
{
#pragma unroll
    for (int u = 0; u < GPW; ++u) {
        const int a = (wave >> 1) + 4 * u;
        if constexpr (REV) {
            const fast::u4 v = __builtin_amdgcn_raw_buffer_load_b128(
                src, (int)(4 * lane_off), 4 * (16 * a + SIDE * (2 * j + (wave & 1))), 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) ck.w[4 * u + i] = v[i];
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                ck.w[4 * u + i] = __builtin_amdgcn_raw_buffer_load_b32(src, (int)(4 * lane_off),
                                                                       4 * (SIDE * (a + 16 * i) + CH * j), 0);
        }
    }
}

template <int PREC, int FMT, bool REV>
__global__ void __launch_bounds__(NT) fft_onchip(OnchipArgs a)
{
    typedef typename Elem<FMT>::raw raw;
    static_assert(sizeof(raw) == 4, "16-bit complex words");
    __shared__ raw ylds[ROWS * YP];        // 80 KB: the rows of one round
    __shared__ float2 ex[SIDE * CH];       // 64 KB: exchange between radix-16 stages
    __shared__ float2 stw[SIDE];           // W_256^e, scaled
    __shared__ float2 twl[2 * SIDE];       // W_65536 lo (scaled) and hi halves

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int w0 = wave & 1, w2 = wave >> 1;
    const int g = lane >> 4, c = lane & 15;
    const int t = 16 * w0 + c;                                // the lane's line in a stage
    // phase-1 load offset, lane part (natural order: row 64 g, column t; digit-reversed: word
    // 4 g of stored row 16 c)
    const unsigned lane_off = REV ? (unsigned)(4 * g + 16 * SIDE * c) : (unsigned)(SIDE * 64 * g + t);
    // exchange-buffer bases: stage-1 writes for even / odd u, stage-2 reads
    // (the odd base is taken at u = 1 and moved back by u = 1's offset, 64 rows)
    const int exw[2] = {ephys(w2, g, t), ephys(w2 + 4, g, t) - 64 * CH};
    const int exr = ephys(4 * g, w2, t);
    // Y bases: writes (row f + 16 g, n2 = 32 j + t), reads (row 32 hb + t, n2 = a2 + 16 i + 64 g)
    const int yw = 16 * g * YP + c;
    const int yr = c * YP + 80 * g;
    const float fac = a.fac, inv = 1.0f / a.fac;

    Ops<PREC> op;
    // permuted operand rows: output row m of a group holds frequency 4 (m % 4) + m / 4
    const int m = lane & 15;
    op.load(reinterpret_cast<const DftOps *>(a.dft_ops), (lane & 48) | (4 * (m & 3) + (m >> 2)));

    // one line per workgroup (grid = lines): a loop over lines would make every lane address
    // loop-invariant, and hoisting them all out of it exhausts the registers
    const long long b = blockIdx.x;
    const __amdgpu_buffer_rsrc_t src = line_rsrc(reinterpret_cast<const raw *>(a.in) + b * a.in_dist);
    const __amdgpu_buffer_rsrc_t dst = line_rsrc(reinterpret_cast<raw *>(a.out) + b * a.out_dist);
    // the tables are read before the first chunks are requested and written to LDS after,
    // so their wait does not cover the chunks' loads
    const bool tab = threadIdx.x < SIDE;
    float2 t0, t1, t2;
    if (tab) {
        t0 = a.stage_tw_s[threadIdx.x];
        t1 = a.tw_lo_s[threadIdx.x];
        t2 = a.tw_hi[threadIdx.x];
    }
    Chunk<PREC, FMT> buf[DEPTH + 1];
#pragma unroll
    for (int j = 0; j < DEPTH; ++j) load_chunk<PREC, FMT, REV>(buf[j], src, j, wave, lane_off);
    if (tab) {
        stw[threadIdx.x] = t0;
        twl[threadIdx.x] = t1;
        twl[SIDE + threadIdx.x] = t2;
    }
    // rounds 1..3 of Y, held: hold[j][u][r - 1] is point (k1 = f + 16 g + 64 r, n2 = 32 j + t)
    raw hold[NCH][GPW][ROUNDS - 1];

    // ---------------- phase 1: columns
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
        if (j + DEPTH < NCH) {
            // issued before this chunk's stage 1 reads its buffer: buf[(j + DEPTH) % (DEPTH
            // + 1)] was consumed by chunk j - 1
            load_chunk<PREC, FMT, REV>(buf[(j + DEPTH) % (DEPTH + 1)], src, j + DEPTH, wave, lane_off);
        }
        __syncthreads();  // the previous chunk's stage 2 has read the exchange buffer
        stage1<PREC, FMT>(buf[j % (DEPTH + 1)].w, op, stw, ex, w2, g, exw);
        // four-step twiddle W^(n2 k1), k1 = f + 16 g + 64 i: one lookup for i = 0 and
        // the step W^(64 n2) for the rest
        const unsigned n2 = (unsigned)(CH * j + t);
        float2 st = tw65536(twl, 64u * n2);
        st.x *= inv;
        st.y *= inv;
#pragma unroll
        for (int u = 0; u < GPW; ++u) {
            const int f = w2 + 4 * u;
            f4 ar, ai;
            stage2<PREC>(ex, op, u, exr, ar, ai);
            float2 w = tw65536(twl, n2 * (unsigned)(f + 16 * g));
            // Y word (row f + 16 g, n2): base + f YP + 32 j + 16 w0 + 16 (j >> 1)
            const int yo = f * YP + CH * j + 16 * w0 + 16 * (j >> 1);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float2 y = cmulf(make_float2(ar[i], ai[i]), w);
                if (i == 0)
                    ylds[yw + yo] = Elem<FMT>::pack(y);
                else
                    hold[j][u][i - 1] = Elem<FMT>::pack(y);
                if (i < 3) w = cmulf(w, st);
            }
        }
    }

    // ---------------- phase 2: rows, four rounds of 64
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        if (r > 0) {
            // the previous round's stage-1 reads of these rows ended at the barrier
            // before its last exchange write
#pragma unroll
            for (int j = 0; j < NCH; ++j)
#pragma unroll
                for (int u = 0; u < GPW; ++u)
                    ylds[yw + (w2 + 4 * u) * YP + CH * j + 16 * w0 + 16 * (j >> 1)] = hold[j][u][r - 1];
        }
        __syncthreads();
#pragma unroll
        for (int hb = 0; hb < 2; ++hb) {
            // rows 32 hb + t of the round; stage-1 operands: points a2 + 16 (4g + i)
            raw w[GPW * 4];
#pragma unroll
            for (int u = 0; u < GPW; ++u)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    w[4 * u + i] = ylds[yr + (CH * hb + 16 * w0) * YP + w2 + 4 * u + 16 * i];
            __syncthreads();  // the previous block's stage 2 has read the exchange buffer
            stage1<PREC, FMT>(w, op, stw, ex, w2, g, exw);
            // X[k1 + 256 k2], k1 = 64 r + 32 hb + t, k2 = f + 16 (4i + g): 16 lanes store 16
            // consecutive k1; uniform part 64 r + 32 hb + 256 (f + 64 i), lane part t + 4096 g
            const unsigned st_off = (unsigned)(t + SIDE * 16 * g);
#pragma unroll
            for (int u = 0; u < GPW; ++u) {
                const int f = w2 + 4 * u;
                f4 ar, ai;
                stage2<PREC>(ex, op, u, exr, ar, ai);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    __builtin_amdgcn_raw_buffer_store_b32(Elem<FMT>::pack(make_float2(ar[i] * fac, ai[i] * fac)), dst,
                                                          (int)(4 * st_off),
                                                          4 * (ROWS * r + CH * hb + SIDE * (f + 64 * i)), 0);
            }
        }
    }
}

}  // namespace onchip

// ------------------------------------------------------------------ strided 2048-point columns
// 32 adjacent columns of 2048 points per workgroup (the column axis of a 2048-row 2D
// transform in one pass; the reference's 2D column kernels, tcfft_half_2d.cu:265-381, take
// these lengths in several strided merges). n = n_lo + 128 n_hi:
//   phase 1: for each (column, n_lo): DFT16 over n_hi (one MFMA stage, exact fp16 inputs),
//            times W_2048^(n_lo k_hi) -> Y[k_hi][n_lo][column];
//   phase 2: for each (column, k_hi): DFT128 over n_lo: a radix-16 MFMA stage over the high
//            digit (n_lo = a + 8 b), times W_128^(a k_b), then a radix-8 VALU stage over a
//            -> X[k_hi + 16 (k_b + 16 k_a)].
// As in the 65536-point kernel the intermediate stays on chip: four rounds of 4 k_hi values,
// round 0 written to LDS by phase 1, rounds 1-3 held in registers (96 words per lane) and
// dumped into LDS at the round's start. Phase 1 needs no exchange buffer and no barrier.
namespace onchip_cols {
using fast::f4;
using fast::h8;
using fast::b8;
using fast::Ops;
using fast::dft16;
using fast::cmulf;
using fast::Elem;

constexpr int NT = 512;
constexpr int LC = kOnchipColsLines;  // 32 columns
constexpr int M = 128;                // n = n_lo + 128 n_hi
constexpr int NCH = 8;                // phase-1 chunks of 16 n_lo values
constexpr int GPW = 4;

// Y word (khl, n_lo, column) of a round (k_hi = 4 r + khl): [khl][n_lo][column], the column
// XOR-swizzled by 16 s, s = bit 0 of (khl ^ n_lo >> 5): phase-1 writes (16 consecutive columns
// x 4 khl) and phase-2 reads (16 consecutive columns x 4 n_lo 32 apart) hit 32 distinct words
// per half wave.
__host__ __device__ constexpr int ycol(int khl, int nlo, int col)
{
    return (khl * M + nlo) * LC + (col ^ (((khl ^ (nlo >> 5)) & 1) << 4));
}
// The rows kind keeps Y as [khl][line][n_lo] with n_lo XOR-swizzled by (line & 15) | 16 (khl & 1):
// its phase-1 writes walk 16 consecutive n_lo (x 4 khl), its phase-2 reads 16 consecutive
// lines (x 4 n_lo 32 apart), both on 32 distinct words per half wave.
template <bool ROWS>
__host__ __device__ constexpr int yaddr(int khl, int nlo, int line)
{
    if constexpr (ROWS)
        return (khl * LC + line) * M + (nlo ^ ((line & 15) | ((khl & 1) << 4)));
    else
        return ycol(khl, nlo, line);
}

// Exchange buffer (phase 2, two khl per block): float2 at [khl & 1][a][k_b][column], the column
// XOR-swizzled by bit 0 of k_b: stage-A writes (16 columns x 4 consecutive k_b) and stage-B
// reads (32 consecutive columns at one k_b) are both conflict-free.
__host__ __device__ constexpr int ecol(int kh2, int a, int kb, int col)
{
    return ((kh2 * 8 + a) * 16 + kb) * LC + (col ^ ((kb & 1) << 4));
}

template <int PREC, int FMT, bool ROWS>
__global__ void __launch_bounds__(NT) fft_onchip_cols(OnchipColsArgs a)
{
    typedef typename Elem<FMT>::raw raw;
    static_assert(sizeof(raw) == 4, "16-bit complex words");
    __shared__ raw ylds[4 * M * LC];     // 64 KB: the rows of one round
    __shared__ float2 ex[2 * 8 * 16 * LC];  // 64 KB: phase-2 exchange (two khl)
    __shared__ float2 twl[kOnchipColsN];    // 16 KB: W_2048 (scaled)
    __shared__ float2 stw[M];               // W_128 (scaled)

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int w0 = wave & 1, w2 = wave >> 1;
    const int g = lane >> 4, c = lane & 15;
    const int col = 16 * w0 + c;  // the lane's column in the MFMA stages

    Ops<PREC> op;
    const int m = lane & 15;  // permuted operand rows: output row 4g + i holds frequency 4i + g
    op.load(reinterpret_cast<const DftOps *>(a.dft_ops), (lane & 48) | (4 * (m & 3) + (m >> 2)));

    const long long img = blockIdx.x / a.col_blocks;
    const int cb = (int)(blockIdx.x - img * a.col_blocks);
    // the block's first word: columns cb * 32 of the image; rows 32 cb of the transform (input)
    // and, stored transposed, word 32 cb (output)
    const raw *src = reinterpret_cast<const raw *>(a.in) + img * a.in_dist + (long long)LC * cb * (ROWS ? kOnchipColsN : 1);
    raw *dst = reinterpret_cast<raw *>(a.out) + img * a.out_dist + (long long)LC * cb;
    // one resource per block: 2048 rows of the pitch (the last line ends 32 words after the
    // last row's start); the rows kind reads 32 contiguous lines
    const int ps_out = ROWS ? a.ps_out : a.ps;
    const int span_in = ROWS ? 4 * LC * kOnchipColsN : 4 * ((kOnchipColsN - 1) * a.ps + LC);
    const int span_out = 4 * ((kOnchipColsN - 1) * ps_out + LC);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<raw *>(src), 0, span_in, 0x00020000);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(dst, 0, span_out, 0x00020000);

    // phase-1 group u of chunk j and the lane's (line, n_lo); the lane loads n_hi = 4 g + i.
    //   columns: line = col, n_lo = 16 j + w2 + 4 u; word line + (n_lo + 128 (4 g + i)) ps: lane
    //            part col + 512 g ps, uniform part (n_lo + 128 i) ps (16 lanes: 64 bytes of a row);
    //   rows:    line = 4 j + u, n_lo = 16 wave + c; word 2048 line + n_lo + 128 (4 g + i): lane
    //            part c + 512 g, uniform part 2048 line + 16 wave + 128 i (a chunk reads 4 whole
    //            lines)
    auto p1_line = [&](int j, int u) { return ROWS ? 4 * j + u : col; };
    auto p1_nlo = [&](int j, int u) { return ROWS ? 16 * wave + c : 16 * j + w2 + 4 * u; };
    const int lane_ld = ROWS ? 4 * (c + 4 * M * g) : 4 * (col + 4 * M * g * a.ps);
    auto load_chunk = [&](raw *w, int j) {
#pragma unroll
        for (int u = 0; u < GPW; ++u)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int uni = ROWS ? 4 * (kOnchipColsN * (4 * j + u) + 16 * wave + M * i)
                                     : 4 * (16 * j + w2 + 4 * u + M * i) * a.ps;
                w[4 * u + i] = __builtin_amdgcn_raw_buffer_load_b32(rs, lane_ld, uni, 0);
            }
    };
    raw buf[2][GPW * 4];
    // the tables are read before the first chunk is requested and written to LDS after it
    float2 tt[kOnchipColsN / NT];
#pragma unroll
    for (int k = 0; k < kOnchipColsN / NT; ++k) tt[k] = a.tw_s[threadIdx.x + NT * k];
    float2 ts = make_float2(0.f, 0.f);
    if (threadIdx.x < M) ts = a.stw_s[threadIdx.x];
    load_chunk(buf[0], 0);
#pragma unroll
    for (int k = 0; k < kOnchipColsN / NT; ++k) twl[threadIdx.x + NT * k] = tt[k];
    if (threadIdx.x < M) stw[threadIdx.x] = ts;
    __syncthreads();

    raw hold[NCH][GPW][3];
    // ---------------- phase 1: one radix-16 stage, twiddle, round 0 to LDS, rounds 1-3 held
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
        if (j + 1 < NCH) load_chunk(buf[(j + 1) & 1], j + 1);
        const raw *w = buf[j & 1];
#pragma unroll
        for (int u = 0; u < GPW; ++u) {
            fast::u4 v = {w[4 * u], w[4 * u + 1], w[4 * u + 2], w[4 * u + 3]};
            const float2 x[4] = {};
            f4 ar, ai;
            dft16<PREC, true>(x, __builtin_bit_cast(h8, v), __builtin_bit_cast(b8, v), op, ar, ai);
            const int nlo = p1_nlo(j, u);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // k_hi = 4 i + g: round i, khl = g
                const float2 y = cmulf(make_float2(ar[i], ai[i]), twl[nlo * (4 * i + g)]);
                if (i == 0)
                    ylds[yaddr<ROWS>(g, nlo, p1_line(j, u))] = Elem<FMT>::pack(y);
                else
                    hold[j][u][i - 1] = Elem<FMT>::pack(y);
            }
        }
    }

    // ---------------- phase 2: four rounds (k_hi = 4 r + khl), two blocks of two khl each
    const int tcol = threadIdx.x & 31, tkb = (threadIdx.x >> 5) & 15;  // stage-B thread map
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (r > 0) {
            // the previous round's stage-A reads of these words ended at the barrier before its
            // last exchange write
#pragma unroll
            for (int j = 0; j < NCH; ++j)
#pragma unroll
                for (int u = 0; u < GPW; ++u) ylds[yaddr<ROWS>(g, p1_nlo(j, u), p1_line(j, u))] = hold[j][u][r - 1];
        }
        __syncthreads();
#pragma unroll
        for (int hb = 0; hb < 2; ++hb) {
            // stage A: group u -> (khl = 2 hb + (p >> 3), a = p & 7), p = w2 + 4 u; lane inputs
            // b = 4 g + i (n_lo = a + 8 b) of column col
            f4 ar[GPW], ai[GPW];
#pragma unroll
            for (int u = 0; u < GPW; ++u) {
                const int p = w2 + 4 * u, khl = 2 * hb + (p >> 3), aa = p & 7;
                fast::u4 v;
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] = ylds[yaddr<ROWS>(khl, aa + 8 * (4 * g + i), col)];
                const float2 x[4] = {};
                dft16<PREC, true>(x, __builtin_bit_cast(h8, v), __builtin_bit_cast(b8, v), op, ar[u], ai[u]);
            }
            __syncthreads();  // the previous block's stage B has read the exchange buffer
#pragma unroll
            for (int u = 0; u < GPW; ++u) {
                const int p = w2 + 4 * u, kh2 = p >> 3, aa = p & 7;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int kb = 4 * i + g;
                    ex[ecol(kh2, aa, kb, col)] = cmulf(make_float2(ar[u][i], ai[u][i]), stw[aa * kb]);
                }
            }
            __syncthreads();
            // stage B: thread -> (line tcol, k_b = tkb), both khl of the block; radix 8 over a,
            // stored at k = k_hi + 16 (k_b + 16 k_a): 32 consecutive lines per output row
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                float2 v[8];
#pragma unroll
                for (int aa = 0; aa < 8; ++aa) v[aa] = ex[ecol(q, aa, tkb, tcol)];
                fast::dft_tail<8>(v, a.sign, a.w16);
                const int khi = 4 * r + 2 * hb + q;
#pragma unroll
                for (int ka = 0; ka < 8; ++ka) {
                    const int k = khi + 16 * (tkb + 16 * ka);
                    const float2 y = make_float2(v[ka].x * a.tail_mul, v[ka].y * a.tail_mul);
                    __builtin_amdgcn_raw_buffer_store_b32(Elem<FMT>::pack(y), rd, 4 * tcol, 4 * k * ps_out,
                                                          kStoreStream);
                }
            }
        }
    }
}

}  // namespace onchip_cols

const void *find_onchip_kernel(int prec, int fmt, bool rev)
{
    using namespace onchip;
    if (prec == PREC_SPLIT && fmt == FMT_C16)
        return rev ? (const void *)fft_onchip<PREC_SPLIT, FMT_C16, true> : (const void *)fft_onchip<PREC_SPLIT, FMT_C16, false>;
    if (prec == PREC_FP16 && fmt == FMT_C16)
        return rev ? (const void *)fft_onchip<PREC_FP16, FMT_C16, true> : (const void *)fft_onchip<PREC_FP16, FMT_C16, false>;
    if (prec == PREC_BF16 && fmt == FMT_CBF16 && !rev) return (const void *)fft_onchip<PREC_BF16, FMT_CBF16, false>;
    return nullptr;
}

const void *find_onchip_cols_kernel(int prec, int fmt, bool rows)
{
    using namespace onchip_cols;
    if (prec == PREC_SPLIT && fmt == FMT_C16)
        return rows ? (const void *)fft_onchip_cols<PREC_SPLIT, FMT_C16, true>
                    : (const void *)fft_onchip_cols<PREC_SPLIT, FMT_C16, false>;
    if (prec == PREC_FP16 && fmt == FMT_C16)
        return rows ? (const void *)fft_onchip_cols<PREC_FP16, FMT_C16, true>
                    : (const void *)fft_onchip_cols<PREC_FP16, FMT_C16, false>;
    if (prec == PREC_BF16 && fmt == FMT_CBF16)
        return rows ? (const void *)fft_onchip_cols<PREC_BF16, FMT_CBF16, true>
                    : (const void *)fft_onchip_cols<PREC_BF16, FMT_CBF16, false>;
    return nullptr;
}

}  // namespace mfft
