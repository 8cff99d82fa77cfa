// Implementation of the fast tile kernels (included by one .hip file per precision).
#pragma once
#include "mfft_fast.h"

namespace mfft {
namespace fast {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

enum { MAP_POINTS = 0, MAP_LINES = 1 };

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ------------------------------------------------------------------ element formats
template <int FMT>
struct Elem;
template <>
struct Elem<FMT_C16> {
    typedef uint32_t raw;
    static __device__ __forceinline__ float2 f(raw w)
    {
        _Float16 re = __builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
        _Float16 im = __builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
        return make_float2((float)re, (float)im);
    }
    static __device__ __forceinline__ raw pack(float2 v) { return pack_c16(v.x, v.y); }
};
template <>
struct Elem<FMT_CBF16> {
    typedef uint32_t raw;
    static __device__ __forceinline__ float2 f(raw w)
    {
        return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
    }
    static __device__ __forceinline__ raw pack(float2 v) { return pack_cbf16(v.x, v.y); }
};
template <>
struct Elem<FMT_C32> {
    typedef float2 raw;
    static __device__ __forceinline__ float2 f(raw w) { return w; }
    static __device__ __forceinline__ raw pack(float2 v) { return v; }
};

// ------------------------------------------------------------------ geometry
// Waves per workgroup (fast_waves_of). A tile is T lines of R points (T = kMaxTile / R;
// cols_lines(R) for the strided kind).
template <int PREC, int KIND, int R>
constexpr int fast_waves()
{
    return fast_waves_of(PREC, R, KIND);
}

template <int R, int KIND, int NW_ = 4>
struct Geo {
    static constexpr int NW = NW_;            // waves per workgroup
    static constexpr int NT = NW_ * 64;       // threads per workgroup
    static constexpr int R16 = R / 16;
    static constexpr int T = KIND == KIND_COLS ? cols_lines(R) : (kMaxTile / R);
    // ROWS*: [line][point], padded points and an odd line pitch (lanes walk either the
    // points of a line or, in the transposed last stage, the lines);
    // COLS: [point][T lines] (every access walks consecutive lines of one point).
    static constexpr int LP = R + R / 16 + 1;
    static constexpr int SIZE = KIND == KIND_COLS ? R * T : T * LP;
    static __device__ __forceinline__ int phys(int t, int p)
    {
        if constexpr (KIND == KIND_COLS)
            return p * T + t;
        else
            return t * LP + p + (p >> 4);
    }
    // stage twiddles staged in LDS (R entries, 32 KB at R = 4096): a read from global
    // memory inside the tile loop would make the wave wait for the next tile's prefetch
    static constexpr int STW = R;
    // four-step table entries staged in LDS per half (lo at [0, TWC), hi at [TWC, 2 TWC))
    static constexpr int TWC = KIND == KIND_COLS ? kFastTwCols : kFastTw;
    static constexpr int PAIRS = R * T / 256;               // MFMA column groups per stage
    static constexpr int MAXP = (PAIRS + NW - 1) / NW;
};

// slot (16 per MFMA group) -> (line t, column col) for a lane mapping
template <int R, int T, int MAP>
__device__ __forceinline__ void slot_tc(int slot, int &t, int &col)
{
    if constexpr (MAP == MAP_POINTS) {
        t = slot / (R / 16);
        col = slot % (R / 16);
    } else {
        t = slot % T;
        col = slot / T;
    }
}

struct Origin {
    int l0b, l1, nv;
    long long in_base, out_base;
};

// Row tiles of 256-point lines: on the matrix cores MFMA group P of a stage is line P in
// both radix-16 stages (16 slots of a line per group) and P = wave + u NW; in the fp32 path
// thread i holds a column of line i / 16 in both stages. Either way every line's LDS
// exchange, and the next tile's overwrite of it, stays inside one wave. LDS operations of a
// wave execute in order, so such a tile needs no workgroup barrier between its stages or
// before the next tile, only a point the compiler does not move LDS accesses across.
template <int PREC, int R, int KIND>
constexpr bool wave_lines() { return KIND == KIND_ROWS && R == 256; }

template <int PREC, int R, int KIND>
__device__ __forceinline__ void tile_sync()
{
    if constexpr (wave_lines<PREC, R, KIND>()) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// The tiles of one workgroup form the progression tile0, tile0 + step, ... The walker keeps
// the current tile as (t0, l1, l2) with tile = t0 + tiles0 (l1 + c1 l2), the step split the
// same way, and the element offsets of the tile's first line; advancing adds the parts with
// at most one carry each, so the loop does no division (the scalar unit is shared by all
// waves of the compute unit, and each wave walks the same tiles).
struct TileWalk {
    int t0, l1, s0, s1, tiles0;
    long long l2, s2;
    long long in_base, out_base;
    long long in_a, in_b, in_c, out_a, out_b, out_c;

    __device__ __forceinline__ void init(const TileArgs &a, int T, long long tile, long long step)
    {
        tiles0 = (a.c0 + T - 1) / T;
        long long r = tile / tiles0;
        t0 = (int)(tile - r * tiles0);
        l1 = (int)(r % a.c1);
        l2 = r / a.c1;
        r = step / tiles0;
        s0 = (int)(step - r * tiles0);
        s1 = (int)(r % a.c1);
        s2 = r / a.c1;
        const long long ti = (long long)T * a.in_s0, to = (long long)T * a.out_s0;
        in_base = t0 * ti + l1 * a.in_s1 + l2 * a.in_s2;
        out_base = t0 * to + l1 * a.out_s1 + l2 * a.out_s2;
        in_a = s0 * ti + s1 * a.in_s1 + s2 * a.in_s2;
        out_a = s0 * to + s1 * a.out_s1 + s2 * a.out_s2;
        in_b = a.in_s1 - tiles0 * ti;
        out_b = a.out_s1 - tiles0 * to;
        in_c = a.in_s2 - a.c1 * a.in_s1;
        out_c = a.out_s2 - a.c1 * a.out_s1;
    }
    __device__ __forceinline__ Origin origin(const TileArgs &a, int T) const
    {
        Origin o;
        o.l0b = t0 * T;
        o.l1 = l1;
        o.nv = min(T, a.c0 - o.l0b);
        o.in_base = in_base;
        o.out_base = out_base;
        return o;
    }
    __device__ __forceinline__ void advance(const TileArgs &a)
    {
        t0 += s0;
        l1 += s1;
        l2 += s2;
        in_base += in_a;
        out_base += out_a;
        if (t0 >= tiles0) {
            t0 -= tiles0;
            ++l1;
            in_base += in_b;
            out_base += out_b;
        }
        if (l1 >= a.c1) {
            l1 -= a.c1;
            ++l2;
            in_base += in_c;
            out_base += out_c;
        }
    }
};

// ------------------------------------------------------------------ DFT16 on MFMA
// Per-lane A operands of the DFT16 (loaded once per kernel).
template <int PREC>
struct Ops {
    h8 hr, hi, lr, li;
    b8 br, bi;
    __device__ __forceinline__ void load(const DftOps *ops, int lane)
    {
        if constexpr (PREC == PREC_BF16) {
            br = *reinterpret_cast<const b8 *>(ops->br[lane]);
            bi = *reinterpret_cast<const b8 *>(ops->bi[lane]);
        } else {
            hr = *reinterpret_cast<const h8 *>(ops->hr[lane]);
            hi = *reinterpret_cast<const h8 *>(ops->hi_[lane]);
            if constexpr (PREC == PREC_SPLIT) {
                lr = *reinterpret_cast<const h8 *>(ops->lr[lane]);
                li = *reinterpret_cast<const h8 *>(ops->li[lane]);
            }
        }
    }
};

// x = 4 complex inputs of this lane (fp32), or `raw16` = the same inputs already in
// fp16/bf16 (exact, from a 16-bit buffer). Returns Y[4g+i] of the lane's column (x16).
template <int PREC, bool EXACT16>
__device__ __forceinline__ void dft16(const float2 *x, const h8 &raw16, const b8 &rawb, const Ops<PREC> &op,
                                      f4 &ar, f4 &ai)
{
    ar = f4{0.f, 0.f, 0.f, 0.f};
    ai = ar;
    if constexpr (PREC == PREC_BF16) {
        b8 b;
        if constexpr (EXACT16) {
            b = rawb;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                b[2 * k] = (__bf16)x[k].x;
                b[2 * k + 1] = (__bf16)x[k].y;
            }
        }
        ar = __builtin_amdgcn_mfma_f32_16x16x32_bf16(op.br, b, ar, 0, 0, 0);
        ai = __builtin_amdgcn_mfma_f32_16x16x32_bf16(op.bi, b, ai, 0, 0, 0);
    } else {
        h8 bh, bl;
        if constexpr (EXACT16) {
            bh = raw16;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                pk_h2 hv, lv;
                if constexpr (PREC == PREC_SPLIT) {
                    split_c16(x[k].x, x[k].y, hv, lv);
                    bl[2 * k] = lv.x;
                    bl[2 * k + 1] = lv.y;
                } else {
                    hv = __builtin_convertvector((pk_f2){x[k].x, x[k].y}, pk_h2);
                }
                bh[2 * k] = hv.x;
                bh[2 * k + 1] = hv.y;
            }
        }
        if constexpr (PREC == PREC_SPLIT) {
            ar = __builtin_amdgcn_mfma_f32_16x16x32_f16(op.lr, bh, ar, 0, 0, 0);
            ai = __builtin_amdgcn_mfma_f32_16x16x32_f16(op.li, bh, ai, 0, 0, 0);
            if constexpr (!EXACT16) {
                ar = __builtin_amdgcn_mfma_f32_16x16x32_f16(op.hr, bl, ar, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_16x16x32_f16(op.hi, bl, ai, 0, 0, 0);
            }
        }
        ar = __builtin_amdgcn_mfma_f32_16x16x32_f16(op.hr, bh, ar, 0, 0, 0);
        ai = __builtin_amdgcn_mfma_f32_16x16x32_f16(op.hi, bh, ai, 0, 0, 0);
    }
}

// Wait for every outstanding vector-memory operation of this wave. Called in the last stage
// before the tile's stores: what is then outstanding is the next tile's prefetch (issued a
// whole tile earlier) and the previous tile's stores. Loads and stores share the counter,
// so a wait for the prefetched registers placed after this tile's stores would also wait
// for those stores, and one placed after the following tile's prefetch would wait for that.
__device__ __forceinline__ void wait_prefetch()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt unconstrained
}

// ------------------------------------------------------------------ first-stage loads
// The raw (undecoded) first-stage operands of one tile, held in registers so the next
// tile's loads can be in flight while the current tile is transformed.
template <int PREC, int R, int KIND, int INF>
struct Prefetch {
    typedef Geo<R, KIND, fast_waves<PREC, KIND, R>()> G;
    typedef typename Elem<INF>::raw raw;
    static constexpr int N = PREC == PREC_FP32 ? 16 : 4 * G::MAXP;
    raw w[N];

    // FULL: every tile of the launch has all its lines (the line count is a multiple of
    // the tile's), so the loads need no per-lane masks. REV: the lines are stored in
    // digit-reversed order (TileArgs::rev_input): the lane's four operands of a group,
    // natural points col + R16 (4g + i), sit at the consecutive positions
    // 16 digit_rev(col) + 4g + i.
    template <bool FULL, bool REV = false>
    __device__ __forceinline__ void load(const TileArgs &a, const Origin &o)
    {
        constexpr int MAP = KIND == KIND_COLS ? MAP_LINES : MAP_POINTS;
        constexpr int R16 = G::R16;
        const raw *src = reinterpret_cast<const raw *>(a.in);
        if constexpr (REV) {
            static_assert(PREC != PREC_FP32, "digit-reversed loads are for the matrix-core kernels");
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            const int g = lane >> 4, c = lane & 15;
#pragma unroll
            for (int u = 0; u < G::MAXP; ++u) {
                const int P = wave + u * G::NW;
                int t, col;
                slot_tc<R, G::T, MAP>(P * 16 + c, t, col);
                const bool ok = (G::PAIRS % G::NW == 0 || P < G::PAIRS) && (FULL || t < o.nv);
                const long long pos = 16LL * digit_rev(col, a.in_rev) + 4 * g;
                if constexpr (KIND == KIND_COLS) {
                    // strided lines: each point a row of its own, consecutive lines across lanes
                    const raw *base = src + o.in_base + (long long)t * a.in_s0 + pos * a.in_ps;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if constexpr (INF == FMT_C32)
                            w[4 * u + i] = ok ? base[i * a.in_ps] : make_float2(0.f, 0.f);
                        else
                            w[4 * u + i] = ok ? base[i * a.in_ps] : 0u;
                    }
                } else {
                    // contiguous lines (in_ps = 1): the four points are one 16-byte (C16) or
                    // two 16-byte (C32) loads when the buffer is 16-byte aligned (every line
                    // and position offset is a multiple of four points)
                    const raw *base = src + o.in_base + (long long)t * a.in_s0 + pos;
                    if constexpr (INF == FMT_C32) {
                        typedef float f4v __attribute__((ext_vector_type(4)));
                        if (a.in_vec) {
                            const f4v z = {0.f, 0.f, 0.f, 0.f};
                            const f4v v0 = ok ? reinterpret_cast<const f4v *>(base)[0] : z;
                            const f4v v1 = ok ? reinterpret_cast<const f4v *>(base)[1] : z;
                            w[4 * u + 0] = make_float2(v0.x, v0.y);
                            w[4 * u + 1] = make_float2(v0.z, v0.w);
                            w[4 * u + 2] = make_float2(v1.x, v1.y);
                            w[4 * u + 3] = make_float2(v1.z, v1.w);
                        } else {
#pragma unroll
                            for (int i = 0; i < 4; ++i) w[4 * u + i] = ok ? base[i] : make_float2(0.f, 0.f);
                        }
                    } else {
                        if (a.in_vec) {
                            const u4 v = ok ? *reinterpret_cast<const u4 *>(base) : u4{0u, 0u, 0u, 0u};
#pragma unroll
                            for (int i = 0; i < 4; ++i) w[4 * u + i] = v[i];
                        } else {
#pragma unroll
                            for (int i = 0; i < 4; ++i) w[4 * u + i] = ok ? base[i] : 0u;
                        }
                    }
                }
            }
            return;
        }
        if constexpr (PREC == PREC_FP32) {
            int t, col;
            slot_tc<R, G::T, MAP>(threadIdx.x, t, col);
            const bool ok = FULL || t < o.nv;
            const raw *base = src + o.in_base + (long long)t * a.in_s0 + (long long)col * a.in_ps;
            const long long step = (long long)R16 * a.in_ps;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if constexpr (INF == FMT_C32)
                    w[j] = ok ? base[j * step] : make_float2(0.f, 0.f);
                else
                    w[j] = ok ? base[j * step] : 0u;
            }
        } else {
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            const int g = lane >> 4, c = lane & 15;
#pragma unroll
            for (int u = 0; u < G::MAXP; ++u) {
                const int P = wave + u * G::NW;
                int t, col;
                slot_tc<R, G::T, MAP>(P * 16 + c, t, col);
                const bool ok = (G::PAIRS % G::NW == 0 || P < G::PAIRS) && (FULL || t < o.nv);
                const raw *base = src + o.in_base + (long long)t * a.in_s0 + (long long)(col + 4 * g * R16) * a.in_ps;
                const long long step = (long long)R16 * a.in_ps;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if constexpr (INF == FMT_C32)
                        w[4 * u + i] = ok ? base[i * step] : make_float2(0.f, 0.f);
                    else
                        w[4 * u + i] = ok ? base[i * step] : 0u;
                }
            }
        }
    }
};

// Four-step twiddles W_N^(mult * (pt0 + k * STEP)), k = 0..3, read before any store.
template <int STEP, int TWC, bool SCALED = true>
__device__ __forceinline__ void tw4(const TileArgs &a, const float2 *twl, unsigned mult, int pt0, float2 *w)
{
    const unsigned mask = (1u << a.tw_bits) - 1u;
    // address-space-qualified scalar views: a load through them cannot become a flat load
    typedef __attribute__((address_space(3))) const float lds_f;
    typedef __attribute__((address_space(1))) const float glb_f;
    if (a.tw_lds) {
        const lds_f *t = (const lds_f *)twl;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const unsigned e = mult * (unsigned)(pt0 + k * STEP);
            const unsigned ih = 2 * (TWC + (e >> a.tw_bits)), il = 2 * (e & mask);
            w[k] = cmulf(make_float2(t[ih], t[ih + 1]), make_float2(t[il], t[il + 1]));
        }
    } else {
        const glb_f *th = (const glb_f *)a.tw_hi, *tl = (const glb_f *)(SCALED ? a.tw_lo_s : a.tw_lo);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const unsigned e = mult * (unsigned)(pt0 + k * STEP);
            const unsigned ih = 2 * (e >> a.tw_bits), il = 2 * (e & mask);
            w[k] = cmulf(make_float2(th[ih], th[ih + 1]), make_float2(tl[il], tl[il + 1]));
        }
    }
}

// One four-step twiddle W_N^e, from LDS (the staged tables) or global memory as in tw4;
// SCALED: the stage scale folded into the lo table (the global path picks the table, the
// LDS path holds whichever the kernel staged).
template <int TWC, bool SCALED = false>
__device__ __forceinline__ float2 tw1(const TileArgs &a, const float2 *twl, unsigned e)
{
    const unsigned mask = (1u << a.tw_bits) - 1u;
    typedef __attribute__((address_space(3))) const float lds_f;
    typedef __attribute__((address_space(1))) const float glb_f;
    if (a.tw_lds) {
        const lds_f *t = (const lds_f *)twl;
        const unsigned ih = 2 * (TWC + (e >> a.tw_bits)), il = 2 * (e & mask);
        return cmulf(make_float2(t[ih], t[ih + 1]), make_float2(t[il], t[il + 1]));
    }
    const glb_f *th = (const glb_f *)a.tw_hi, *tl = (const glb_f *)(SCALED ? a.tw_lo_s : a.tw_lo);
    const unsigned ih = 2 * (e >> a.tw_bits), il = 2 * (e & mask);
    return cmulf(make_float2(th[ih], th[ih + 1]), make_float2(tl[il], tl[il + 1]));
}

// Line factor of the four-step twiddle exponent: the line index along the twiddle level,
// digit-reversed (TileArgs::tw_rev) in the first pass of a digit-reversed plan.
template <bool REV>
__device__ __forceinline__ unsigned tw_line(const TileArgs &a, const Origin &o, int t)
{
    const int line = a.tw_level ? o.l1 : o.l0b + t;
    if constexpr (REV) return (unsigned)digit_rev(line, a.tw_rev);
    return (unsigned)line;
}

// ------------------------------------------------------------------ radix-16 stage (MFMA)
// FIRST: operands are the prefetched raw inputs; LAST: results go to global memory,
// multiplied by the four-step twiddles when TW (a.tw_N != 0), else by a.last_mul.
// stw: stage twiddles (LDS or global, scale folded in); twl: the four-step tables staged
// in LDS (lo at [0, TWC), hi at [TWC, 2 TWC)) when a.tw_lds, else they are read
// from global memory (a.tw_lo_s, a.tw_hi). The two sources are separate code paths so
// each keeps its address space (LDS reads wait on the LDS counter only, never on the
// next tile's prefetch loads).
template <int PREC, int R, int KIND, int MAP, bool FIRST, bool LAST, int INF, int OUTF, bool TW = false,
          bool REV = false>
__device__ __forceinline__ void mstage(float2 *lds, const TileArgs &a, const Origin &o, int s, const Ops<PREC> &op,
                                       const Prefetch<PREC, R, KIND, INF> &pf, const float2 *stw,
                                       const float2 *twl)
{
    typedef Geo<R, KIND, fast_waves<PREC, KIND, R>()> G;
    constexpr int T = G::T, R16 = G::R16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, c = lane & 15;
    constexpr bool EXACT = FIRST && ((INF == FMT_C16 && PREC != PREC_BF16) || (INF == FMT_CBF16 && PREC == PREC_BF16));

    f4 accr[G::MAXP], acci[G::MAXP];
#pragma unroll
    for (int u = 0; u < G::MAXP; ++u) {
        const int P = wave + u * G::NW;
        if (G::PAIRS % G::NW != 0 && P >= G::PAIRS) break;
        int t, col;
        slot_tc<R, T, MAP>(P * 16 + c, t, col);
        float2 x[4];
        h8 r16;
        b8 rb;
        if constexpr (FIRST) {
            if constexpr (EXACT) {
                u4 v = {pf.w[4 * u], pf.w[4 * u + 1], pf.w[4 * u + 2], pf.w[4 * u + 3]};
                r16 = __builtin_bit_cast(h8, v);
                rb = __builtin_bit_cast(b8, v);
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) x[i] = Elem<INF>::f(pf.w[4 * u + i]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) x[i] = lds[G::phys(t, col + (4 * g + i) * R16)];
        }
        dft16<PREC, EXACT>(x, r16, rb, op, accr[u], acci[u]);
    }
    if constexpr (!FIRST && !LAST) __syncthreads();  // all reads of the tile precede the writes
    if constexpr (LAST) wait_prefetch();
    // Last stage with the four-step twiddle: the lane's outputs of group u are the points
    // col + R16 (4g + i), i = 0..3, so their twiddles are w_i = W^(mult (col + 4g R16)) *
    // s^i with s = W^(mult R16): one table lookup for w_0, products for the rest. When the
    // lane's line is the same for every group, s is looked up once per tile.
    constexpr bool SAME_T = MAP == MAP_LINES && (16 * G::NW) % T == 0;
    float2 tw_step = make_float2(0.f, 0.f);
    if constexpr (LAST && TW && SAME_T) {
        int t, col;
        slot_tc<R, T, MAP>(wave * 16 + c, t, col);
        const unsigned mult = tw_line<REV>(a, o, t);
        tw_step = tw1<G::TWC, true>(a, twl, mult * (unsigned)R16);
        const float inv = 1.0f / a.last_mul;  // the table's stage scale, a power of two
        tw_step.x *= inv;
        tw_step.y *= inv;
    }
#pragma unroll
    for (int u = 0; u < G::MAXP; ++u) {
        const int P = wave + u * G::NW;
        if (G::PAIRS % G::NW != 0 && P >= G::PAIRS) break;
        int t, col;
        slot_tc<R, T, MAP>(P * 16 + c, t, col);
        if constexpr (!LAST) {
            const int q = col % s, p = col / s;
            float2 w[4];  // all twiddles first: the reads do not wait behind the tile writes
#pragma unroll
            for (int i = 0; i < 4; ++i) w[i] = stw[p * (4 * g + i) * s];  // W_R^(p*tq*s), scale folded in
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int tq = 4 * g + i;
                lds[G::phys(t, q + s * (16 * p + tq))] = cmulf(make_float2(accr[u][i], acci[u][i]), w[i]);
            }
        } else {
            if (t < o.nv) {
                typedef typename Elem<OUTF>::raw raw;
                raw *dst = reinterpret_cast<raw *>(a.out) + o.out_base + (long long)t * a.out_s0;
                float2 w[4];
                if constexpr (TW) {
                    const unsigned mult = tw_line<REV>(a, o, t);
                    float2 st = tw_step;
                    if constexpr (!SAME_T) {
                        st = tw1<G::TWC, true>(a, twl, mult * (unsigned)R16);
                        const float inv = 1.0f / a.last_mul;
                        st.x *= inv;
                        st.y *= inv;
                    }
                    w[0] = tw1<G::TWC, true>(a, twl, mult * (unsigned)(col + R16 * 4 * g));
                    w[1] = cmulf(w[0], st);
                    w[2] = cmulf(w[1], st);
                    w[3] = cmulf(w[2], st);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int pt = col + R16 * (4 * g + i);
                    float2 y = make_float2(accr[u][i], acci[u][i]);
                    if constexpr (TW) {
                        y = cmulf(y, w[i]);
                    } else {
                        y.x *= a.last_mul;
                        y.y *= a.last_mul;
                    }
                    dst[(long long)pt * a.out_ps] = Elem<OUTF>::pack(y);
                }
            }
        }
    }
    if constexpr (!LAST) tile_sync<PREC, R, KIND>();
}

// ------------------------------------------------------------------ radix-16 stage (fp32 VALU)
__device__ __forceinline__ void dft4v(float2 *v, int sign)
{
    float2 s0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    float2 d0 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    float2 s1 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    float2 d1 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    float2 r = make_float2(-sign * d1.y, sign * d1.x);
    v[0] = make_float2(s0.x + s1.x, s0.y + s1.y);
    v[1] = make_float2(d0.x + r.x, d0.y + r.y);
    v[2] = make_float2(s0.x - s1.x, s0.y - s1.y);
    v[3] = make_float2(d0.x - r.x, d0.y - r.y);
}

__device__ __forceinline__ void dft16v(float2 *v, int sign, const float *w16)
{
    float2 a[4][4];
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) {
        float2 t[4] = {v[j2], v[4 + j2], v[8 + j2], v[12 + j2]};
        dft4v(t, sign);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) a[k1][j2] = t[k1];
    }
    const float c1 = w16[0], s1 = w16[1], h = w16[2];
    const float wr[10] = {1.f, c1, h, s1, 0.f, -s1, -h, -c1, -1.f, -c1};
    const float wi[10] = {0.f, s1, h, c1, 1.f, c1, h, s1, 0.f, -s1};
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1)
#pragma unroll
        for (int j2 = 1; j2 < 4; ++j2) {
            const int e = k1 * j2;
            a[k1][j2] = cmulf(a[k1][j2], make_float2(wr[e], sign * wi[e]));
        }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        dft4v(a[k1], sign);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = a[k1][k2];
    }
}

// One thread = one 16-point column per stage (tile of 4096 points / 256 threads).
template <int R, int KIND, int MAP, bool FIRST, bool LAST, int INF, int OUTF, bool TW = false>
__device__ __forceinline__ void vstage16(float2 *lds, const TileArgs &a, const Origin &o, int s,
                                         const Prefetch<PREC_FP32, R, KIND, INF> &pf, const float2 *stw,
                                         const float2 *twl)
{
    typedef Geo<R, KIND, fast_waves<PREC_FP32, KIND, R>()> G;
    constexpr int T = G::T, R16 = G::R16;
    static_assert(R * G::T / 16 == G::NT, "one column per thread");
    int t, col;
    slot_tc<R, T, MAP>(threadIdx.x, t, col);
    float2 v[16];
    if constexpr (FIRST) {
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = Elem<INF>::f(pf.w[j]);
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = lds[G::phys(t, col + j * R16)];
        if constexpr (!LAST) tile_sync<PREC_FP32, R, KIND>();
    }
    dft16v(v, a.sign, a.w16);
    if constexpr (!LAST) {
        const int q = col % s, p = col / s;
#pragma unroll
        for (int tq = 0; tq < 16; ++tq)
            lds[G::phys(t, q + s * (16 * p + tq))] = cmulf(v[tq], stw[p * tq * s]);
        tile_sync<PREC_FP32, R, KIND>();
    } else {
        wait_prefetch();
        if (t >= o.nv) return;
        typedef typename Elem<OUTF>::raw raw;
        raw *dst = reinterpret_cast<raw *>(a.out) + o.out_base + (long long)t * a.out_s0;
        const unsigned mult = (unsigned)(a.tw_level ? o.l1 : o.l0b + t);
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
            float2 w[4];
            if constexpr (TW) tw4<4 * R16, G::TWC>(a, twl, mult, col + R16 * q4, w);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int tq = q4 + 4 * k;
                const int pt = col + R16 * tq;
                float2 y = v[tq];
                if constexpr (TW) {
                    y = cmulf(y, w[k]);
                } else {
                    y.x *= a.last_mul;
                    y.y *= a.last_mul;
                }
                dst[(long long)pt * a.out_ps] = Elem<OUTF>::pack(y);
            }
        }
    }
}

// ------------------------------------------------------------------ tail stage (radix 2/4/8, VALU)
template <int RADIX>
__device__ __forceinline__ void dft_tail(float2 *v, int sign, const float *w16)
{
    if constexpr (RADIX == 2) {
        float2 x0 = v[0], x1 = v[1];
        v[0] = make_float2(x0.x + x1.x, x0.y + x1.y);
        v[1] = make_float2(x0.x - x1.x, x0.y - x1.y);
    } else if constexpr (RADIX == 4) {
        dft4v(v, sign);
    } else {
        const float h = w16[2];
        float2 a4[4], b4[4];
#pragma unroll
        for (int j2 = 0; j2 < 4; ++j2) {
            float2 x0 = v[j2], x1 = v[4 + j2];
            a4[j2] = make_float2(x0.x + x1.x, x0.y + x1.y);
            b4[j2] = make_float2(x0.x - x1.x, x0.y - x1.y);
        }
        b4[1] = make_float2(h * (b4[1].x - sign * b4[1].y), h * (b4[1].y + sign * b4[1].x));
        b4[2] = make_float2(-sign * b4[2].y, sign * b4[2].x);
        b4[3] = make_float2(h * (-b4[3].x - sign * b4[3].y), h * (-b4[3].y + sign * b4[3].x));
        dft4v(a4, sign);
        dft4v(b4, sign);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            v[2 * k2] = a4[k2];
            v[2 * k2 + 1] = b4[k2];
        }
    }
}

// stw: the stage twiddles with the matrix-core scale a.last_mul folded in (the table the
// radix-16 stages use, staged in LDS); this VALU stage scales by `mul` instead, so its
// products take mul / last_mul (a power of two) on the input side.
template <int R, int KIND, int RADIX, int NW>
__device__ __forceinline__ void tstage(float2 *lds, const TileArgs &a, int s, float mul, const float2 *stw)
{
    typedef Geo<R, KIND, NW> G;
    constexpr int T = G::T;
    constexpr int COLS = R / RADIX;
    constexpr int ITEMS = COLS * T;
    constexpr int NT = G::NT;
    constexpr int PER = (ITEMS + NT - 1) / NT;
    float2 v[PER][RADIX];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int I = threadIdx.x + k * NT;
        if (ITEMS % NT != 0 && I >= ITEMS) break;
        const int t = KIND == KIND_COLS ? I % T : I / COLS;
        const int col = KIND == KIND_COLS ? I / T : I % COLS;
#pragma unroll
        for (int j = 0; j < RADIX; ++j) v[k][j] = lds[G::phys(t, col + j * COLS)];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int I = threadIdx.x + k * NT;
        if (ITEMS % NT != 0 && I >= ITEMS) break;
        const int t = KIND == KIND_COLS ? I % T : I / COLS;
        const int col = KIND == KIND_COLS ? I / T : I % COLS;
        dft_tail<RADIX>(v[k], a.sign, a.w16);
        const int q = col % s, p = col / s;
        const float comp = mul / a.last_mul;
#pragma unroll
        for (int tq = 0; tq < RADIX; ++tq) {
            float2 y = make_float2(v[k][tq].x * mul, v[k][tq].y * mul);
            if (p) y = cmulf(make_float2(v[k][tq].x * comp, v[k][tq].y * comp), stw[p * tq * s]);
            lds[G::phys(t, q + s * (RADIX * p + tq))] = y;
        }
    }
    __syncthreads();
}

// The radix-2/4/8 tail as the LAST stage (R = 16 * TAIL: one matrix-core stage, then
// this): reads LDS, and stores straight to global memory with the four-step twiddle (the
// unscaled tables: this stage is VALU, no x16 operand factor), the tail's stage scale and
// the output conversion. Lanes walk the points of a line (ROWS) or the lines (the
// transposed and strided kinds), whichever is contiguous in the output.
template <int R, int KIND, int RADIX, int NW, int OUTF, bool TW, bool REV = false>
__device__ __forceinline__ void tstage_last(float2 *lds, const TileArgs &a, const Origin &o, const float2 *twl)
{
    typedef Geo<R, KIND, NW> G;
    constexpr int T = G::T;
    constexpr int COLS = R / RADIX;  // = 16: the stride of the last stage
    constexpr int ITEMS = COLS * T;
    constexpr int NT = G::NT;
    constexpr int PER = (ITEMS + NT - 1) / NT;
    typedef typename Elem<OUTF>::raw raw;
    wait_prefetch();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int I = threadIdx.x + k * NT;
        if (ITEMS % NT != 0 && I >= ITEMS) break;
        const int t = KIND == KIND_ROWS ? I / COLS : I % T;
        const int col = KIND == KIND_ROWS ? I % COLS : I / T;
        float2 v[RADIX];
#pragma unroll
        for (int j = 0; j < RADIX; ++j) v[j] = lds[G::phys(t, col + j * COLS)];
        if (t >= o.nv) continue;
        dft_tail<RADIX>(v, a.sign, a.w16);
        raw *dst = reinterpret_cast<raw *>(a.out) + o.out_base + (long long)t * a.out_s0;
        const unsigned mult = tw_line<REV>(a, o, t);
#pragma unroll
        for (int q4 = 0; q4 < RADIX; q4 += 4) {
            float2 w[4];
            if constexpr (TW && RADIX >= 4)
                tw4<COLS, G::TWC, false>(a, twl, mult, col + COLS * q4, w);
#pragma unroll
            for (int kk = 0; kk < (RADIX < 4 ? RADIX : 4); ++kk) {
                const int tq = q4 + kk;
                const int pt = col + COLS * tq;
                float2 y = make_float2(v[tq].x * a.tail_mul, v[tq].y * a.tail_mul);
                if constexpr (TW) {
                    if constexpr (RADIX >= 4) {
                        y = cmulf(y, w[kk]);
                    } else {
                        y = cmulf(y, tw1<G::TWC>(a, twl, mult * (unsigned)pt));
                    }
                }
                dst[(long long)pt * a.out_ps] = Elem<OUTF>::pack(y);
            }
        }
    }
}

// ------------------------------------------------------------------ kernel
// The persistent tile loop. TW: the last stage multiplies by the four-step twiddles
// (a.tw_N != 0); FULL: every tile is full (unmasked loads). The loop is instantiated once
// per case so no variant carries another's registers, selects or branches (a masked and
// an unmasked load path in one loop made the compiler wait for every outstanding memory
// operation, the previous tile's stores included, before each tile's loads).
// Stage twiddles and the (small) four-step tables into LDS, the DFT16 operands into
// registers.
template <int PREC, int R, int KIND>
__device__ __forceinline__ void stage_tables(const TileArgs &a, float2 *s_stw, float2 *s_tw, Ops<PREC> &op)
{
    typedef Geo<R, KIND, fast_waves<PREC, KIND, R>()> G;
    constexpr int NS = __builtin_ctz(R) / 4;
    for (int i = threadIdx.x; i < R; i += G::NT) s_stw[i] = a.stage_tw_s[i];
    if (a.tw_N && a.tw_lds) {
        const int nlo = 1 << a.tw_bits, nhi = (a.tw_N + nlo - 1) >> a.tw_bits;
        // the tail-as-last stage (NS == 1) is VALU: it takes the unscaled lo table
        const float2 *lo = NS == 1 ? a.tw_lo : a.tw_lo_s;
        for (int i = threadIdx.x; i < nlo; i += G::NT) s_tw[i] = lo[i];
        for (int i = threadIdx.x; i < nhi; i += G::NT) s_tw[G::TWC + i] = a.tw_hi[i];
    }
    if constexpr (PREC != PREC_FP32) op.load(reinterpret_cast<const DftOps *>(a.dft_ops), threadIdx.x & 63);
}

template <int PREC, int R, int KIND, int INF, int OUTF, bool TW, bool FULL, bool REV = false>
__device__ __forceinline__ void fast_tiles(const TileArgs &a, float2 *lds, float2 *stw, float2 *s_tw)
{
    typedef Geo<R, KIND, fast_waves<PREC, KIND, R>()> G;
    constexpr int L2 = __builtin_ctz(R);
    constexpr int NS = L2 / 4;
    constexpr int TAIL = 1 << (L2 % 4);
    constexpr int MAP_IN = KIND == KIND_COLS ? MAP_LINES : MAP_POINTS;
    constexpr int MAP_OUT = KIND == KIND_ROWS ? MAP_POINTS : MAP_LINES;
    // Each of the 8 XCDs (workgroups are dealt to them round-robin) walks a contiguous
    // range of tiles, so tiles that share cache lines meet in the same L2.
    const long long ntiles = (long long)((a.c0 + G::T - 1) / G::T) * a.c1 * a.c2;
    long long tile, step, end;
    if ((gridDim.x & 7) == 0) {
        const long long per = (ntiles + 7) >> 3;
        const int xcd = blockIdx.x & 7;
        tile = xcd * per + (blockIdx.x >> 3);
        step = gridDim.x >> 3;
        end = min(ntiles, (xcd + 1) * per);
    } else {
        tile = blockIdx.x;
        step = gridDim.x;
        end = ntiles;
    }
    Prefetch<PREC, R, KIND, INF> cur, nxt;
    TileWalk walk;
    walk.init(a, G::T, tile < end ? tile : 0, step);
    Origin o = walk.origin(a, G::T);
    if (tile < end) cur.template load<FULL, REV>(a, o);
    // the tables are loaded after the first tile's operands, so the two global-memory
    // round trips overlap (a workgroup with one tile waits for one, not two)
    Ops<PREC> op;
    stage_tables<PREC, R, KIND>(a, stw, s_tw, op);
    __syncthreads();
    // the loop starts with no loads in flight: otherwise the first tile's pending loads
    // make every iteration wait for all loads (its own prefetch included) before its
    // first product
    wait_prefetch();
    while (tile < end) {
        const long long next = tile + step;
        Origin on = o;
        if (next < end) {
            walk.advance(a);
            on = walk.origin(a, G::T);
            nxt.template load<FULL, REV>(a, on);
        }
        int s = 1;
        if constexpr (PREC == PREC_FP32)
            vstage16<R, KIND, MAP_IN, true, false, INF, OUTF>(lds, a, o, s, cur, stw, s_tw);
        else
            mstage<PREC, R, KIND, MAP_IN, true, false, INF, OUTF>(lds, a, o, s, op, cur, stw, s_tw);
        s = 16;
        if constexpr (NS == 1) {
            // R = 16 * TAIL: the tail is the last stage and stores to global memory
            tstage_last<R, KIND, TAIL, G::NW, OUTF, TW, REV>(lds, a, o, s_tw);
        } else {
            if constexpr (TAIL > 1) {
                tstage<R, KIND, TAIL, G::NW>(lds, a, s, a.tail_mul, stw);
                s *= TAIL;
            }
#pragma unroll
            for (int k = 1; k < NS - 1; ++k) {
                if constexpr (PREC == PREC_FP32)
                    vstage16<R, KIND, MAP_IN, false, false, INF, OUTF>(lds, a, o, s, cur, stw, s_tw);
                else
                    mstage<PREC, R, KIND, MAP_IN, false, false, INF, OUTF>(lds, a, o, s, op, cur, stw, s_tw);
                s *= 16;
            }
            if constexpr (PREC == PREC_FP32)
                vstage16<R, KIND, MAP_OUT, false, true, INF, OUTF, TW>(lds, a, o, s, cur, stw, s_tw);
            else
                mstage<PREC, R, KIND, MAP_OUT, false, true, INF, OUTF, TW, REV>(lds, a, o, s, op, cur, stw, s_tw);
        }
        tile_sync<PREC, R, KIND>();  // the next tile's first stage overwrites LDS
        tile = next;
        o = on;
        cur = nxt;
    }
}

// R = 16^n * TAIL (TAIL in {1,2,4,8}, n >= 2): schedule [16, TAIL?, 16, ..., 16].
// Persistent: workgroup b transforms tiles b, b + gridDim.x, ...; the first-stage
// operands of the next tile are loaded before the current tile is transformed.
template <int PREC, int R, int KIND, int INF, int OUTF>
__global__ void __launch_bounds__((fast_waves<PREC, KIND, R>() * 64)) fft_fast(TileArgs a)
{
    typedef Geo<R, KIND, fast_waves<PREC, KIND, R>()> G;
    __shared__ float2 lds[G::SIZE];
    __shared__ float2 s_stw[G::STW > 0 ? G::STW : 1];
    __shared__ float2 s_tw[2 * G::TWC];
    constexpr int L2 = __builtin_ctz(R);
    constexpr int NS = L2 / 4;
    constexpr int TAIL = 1 << (L2 % 4);
    static_assert(NS >= 2 || (NS == 1 && TAIL > 1), "fast kernels need R >= 32");

    const bool full = a.c0 % G::T == 0;
    if constexpr (fast_rev_ok(PREC, INF)) {
        if (a.rev_input) {
            if (a.tw_N) {
                if (full)
                    fast_tiles<PREC, R, KIND, INF, OUTF, true, true, true>(a, lds, s_stw, s_tw);
                else
                    fast_tiles<PREC, R, KIND, INF, OUTF, true, false, true>(a, lds, s_stw, s_tw);
            } else {
                if (full)
                    fast_tiles<PREC, R, KIND, INF, OUTF, false, true, true>(a, lds, s_stw, s_tw);
                else
                    fast_tiles<PREC, R, KIND, INF, OUTF, false, false, true>(a, lds, s_stw, s_tw);
            }
            return;
        }
    }
    if (a.tw_N) {
        if (full)
            fast_tiles<PREC, R, KIND, INF, OUTF, true, true>(a, lds, s_stw, s_tw);
        else
            fast_tiles<PREC, R, KIND, INF, OUTF, true, false>(a, lds, s_stw, s_tw);
    } else {
        if (full)
            fast_tiles<PREC, R, KIND, INF, OUTF, false, true>(a, lds, s_stw, s_tw);
        else
            fast_tiles<PREC, R, KIND, INF, OUTF, false, false>(a, lds, s_stw, s_tw);
    }
}

template <int PREC, int KIND, int INF, int OUTF>
fast_fn pick_r(int R)
{
    switch (R) {
    case 64: return fft_fast<PREC, 64, KIND, INF, OUTF>;
    case 128: return fft_fast<PREC, 128, KIND, INF, OUTF>;
    case 256: return fft_fast<PREC, 256, KIND, INF, OUTF>;
    default: break;
    }
    switch (R) {
    case 512: return fft_fast<PREC, 512, KIND, INF, OUTF>;
    case 1024: return fft_fast<PREC, 1024, KIND, INF, OUTF>;
    default: break;
    }
    if constexpr (KIND != KIND_COLS) {
        switch (R) {
        case 2048: return fft_fast<PREC, 2048, KIND, INF, OUTF>;
        case 4096: return fft_fast<PREC, 4096, KIND, INF, OUTF>;
        default: break;
        }
    }
    return nullptr;
}

template <int PREC, int INF, int OUTF>
fast_fn pick_kind(int R, int kind)
{
    switch (kind) {
    case KIND_ROWS: return pick_r<PREC, KIND_ROWS, INF, OUTF>(R);
    case KIND_ROWS_TR: return pick_r<PREC, KIND_ROWS_TR, INF, OUTF>(R);
    case KIND_COLS: return pick_r<PREC, KIND_COLS, INF, OUTF>(R);
    default: return nullptr;
    }
}

}  // namespace fast
}  // namespace mfft
