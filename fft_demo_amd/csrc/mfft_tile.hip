// Tile kernel of the MI355X matrix-core FFT.
//
// Method (the reference's, tcfft_half.cu:69-161, 266-390): every radix-16 stage is a
// 16x16 complex DFT-matrix product on the matrix cores, twiddles are elementwise
// multiplies, and the stride permutations go through on-chip shared memory (LDS).
//
// MI355X design: one workgroup (4 wave64s) owns an R-point x T-line tile held in LDS as
// fp32 complex. A radix-16 stage folds real and imaginary parts into the K dimension,
//     [Yr] = [ Fr  -sFi ] [Xr]      K = 32 = 16 complex inputs interleaved (re, im),
//     [Yi]   [ sFi  Fr  ] [Xi]
// so one v_mfma_f32_16x16x32_{f16,bf16} per output half transforms 16 columns. The B
// operand of lane (g = lane/16, c = lane%16) is the interleaved complex inputs 4g..4g+3
// of column c -- exactly the memory order of 4 consecutive complex values -- and the
// accumulator returns outputs 4g..4g+3 of the same column, so stages chain without
// register shuffles. Split precision adds the lo*hi and hi*lo products (3 MFMAs per
// half). Stages follow a self-sorting Stockham (decimation in frequency) schedule so
// natural-order input gives natural-order output.
#include "mfft_common.h"
#include <hip/hip_ext.h>

namespace mfft {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ element I/O
__device__ __forceinline__ float2 cload1(const void *base, long long i, int fmt)
{
    if (fmt == FMT_C32) return reinterpret_cast<const float2 *>(base)[i];
    uint32_t w = reinterpret_cast<const uint32_t *>(base)[i];
    if (fmt == FMT_C16) {
        _Float16 re = __builtin_bit_cast(_Float16, (uint16_t)(w & 0xffff));
        _Float16 im = __builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
        return make_float2((float)re, (float)im);
    }
    return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

__device__ __forceinline__ uint32_t pack2(float re, float im, int fmt)
{
    return fmt == FMT_C16 ? pack_c16(re, im) : pack_cbf16(re, im);
}

__device__ __forceinline__ void cstore1(void *base, long long i, int fmt, float2 v)
{
    if (fmt == FMT_C32) {
        reinterpret_cast<float2 *>(base)[i] = v;
        return;
    }
    reinterpret_cast<uint32_t *>(base)[i] = pack2(v.x, v.y, fmt);
}

// 4 consecutive complex elements starting at a multiple of 4.
__device__ __forceinline__ void cload4(const void *base, long long i, int fmt, float2 v[4])
{
    if (fmt == FMT_C32) {
        const float4 *p = reinterpret_cast<const float4 *>(reinterpret_cast<const float2 *>(base) + i);
        float4 a = p[0], b = p[1];
        v[0] = make_float2(a.x, a.y);
        v[1] = make_float2(a.z, a.w);
        v[2] = make_float2(b.x, b.y);
        v[3] = make_float2(b.z, b.w);
        return;
    }
    uint4 w = *reinterpret_cast<const uint4 *>(reinterpret_cast<const uint32_t *>(base) + i);
    uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (fmt == FMT_C16) {
            _Float16 re = __builtin_bit_cast(_Float16, (uint16_t)(ws[k] & 0xffff));
            _Float16 im = __builtin_bit_cast(_Float16, (uint16_t)(ws[k] >> 16));
            v[k] = make_float2((float)re, (float)im);
        } else {
            v[k] = make_float2(__uint_as_float(ws[k] << 16), __uint_as_float(ws[k] & 0xffff0000u));
        }
    }
}

__device__ __forceinline__ void cstore4(void *base, long long i, int fmt, const float2 v[4])
{
    if (fmt == FMT_C32) {
        float4 *p = reinterpret_cast<float4 *>(reinterpret_cast<float2 *>(base) + i);
        p[0] = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
        p[1] = make_float4(v[2].x, v[2].y, v[3].x, v[3].y);
        return;
    }
    uint4 w;
    w.x = pack2(v[0].x, v[0].y, fmt);
    w.y = pack2(v[1].x, v[1].y, fmt);
    w.z = pack2(v[2].x, v[2].y, fmt);
    w.w = pack2(v[3].x, v[3].y, fmt);
    *reinterpret_cast<uint4 *>(reinterpret_cast<uint32_t *>(base) + i) = w;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ------------------------------------------------------------------ tile geometry
// LDS layout: lines are grouped in blocks of TL interleaved lines; element
// e = block*(R*TL) + point*TL + (line % TL), padded by one slot every 16.
template <int R_, int T_, int TL_>
struct GeoC {
    static constexpr int R = R_, T = T_, TL = TL_;
    static constexpr int MAXP = ((R * T / 16 + 15) / 16 + kWaves - 1) / kWaves;  // MFMA groups per wave
    __device__ __forceinline__ int phys(int t, int point) const
    {
        int b = t / TL, tl = t - b * TL;
        return lds_pad(b * (R * TL) + point * TL + tl);
    }
};

struct GeoR {
    int R, T, TL;
    static constexpr int MAXP = kMaxPairsPerWave;
    __device__ __forceinline__ int phys(int t, int point) const
    {
        int b = t / TL, tl = t - b * TL;
        return lds_pad(b * (R * TL) + point * TL + tl);
    }
};

// ------------------------------------------------------------------ radix-16 on MFMA
template <int PREC>
__device__ __forceinline__ void dft16_mfma(const float2 x[4], const DftOps *ops, int lane, float out_r[4],
                                           float out_i[4])
{
    f4 zr = {0.f, 0.f, 0.f, 0.f};
    f4 ar = zr, ai = zr;
    if constexpr (PREC == PREC_BF16) {
        b8 b;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            b[2 * k] = (__bf16)x[k].x;
            b[2 * k + 1] = (__bf16)x[k].y;
        }
        b8 Ar = *reinterpret_cast<const b8 *>(ops->br[lane]);
        b8 Ai = *reinterpret_cast<const b8 *>(ops->bi[lane]);
        ar = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ar, b, ar, 0, 0, 0);
        ai = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ai, b, ai, 0, 0, 0);
    } else {
        h8 bh, bl;
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
        h8 Ar = *reinterpret_cast<const h8 *>(ops->hr[lane]);
        h8 Ai = *reinterpret_cast<const h8 *>(ops->hi_[lane]);
        if constexpr (PREC == PREC_SPLIT) {
            h8 Lr = *reinterpret_cast<const h8 *>(ops->lr[lane]);
            h8 Li = *reinterpret_cast<const h8 *>(ops->li[lane]);
            ar = __builtin_amdgcn_mfma_f32_16x16x32_f16(Lr, bh, ar, 0, 0, 0);
            ai = __builtin_amdgcn_mfma_f32_16x16x32_f16(Li, bh, ai, 0, 0, 0);
            ar = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ar, bl, ar, 0, 0, 0);
            ai = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ai, bl, ai, 0, 0, 0);
        }
        ar = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ar, bh, ar, 0, 0, 0);
        ai = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ai, bh, ai, 0, 0, 0);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        out_r[k] = ar[k];
        out_i[k] = ai[k];
    }
}

// One radix-16 Stockham stage on the matrix cores, in place in LDS.
// Stage with stride s (product of earlier radices): column col = q + s*p reads points
// col + j*R/16 (j = 0..15) and writes point q + s*(16p + t) scaled by W_{R/s}^{p*t}.
template <int PREC, class Geo>
__device__ void stage16_mfma(float2 *lds, const Geo &G, int s, float mul, const TileArgs &a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, c = lane & 15;
    const int R = G.R, TL = G.TL;
    const int R16 = R >> 4;
    const int slots = R16 * G.T;                 // (line, column) slots of the tile
    const int pairs = (slots + 15) >> 4;         // MFMA column groups
    const int blk = R16 * TL;                    // slots per block of TL interleaved lines
    const DftOps *ops = reinterpret_cast<const DftOps *>(a.dft_ops);
    const float m16 = mul * (1.0f / 16.0f);      // operands carry a factor 16

    constexpr int MAXP = Geo::MAXP;
    float2 res[MAXP][4];
    int dst[MAXP][4];
    bool ok[MAXP];
#pragma unroll
    for (int u = 0; u < MAXP; ++u) ok[u] = false;
#pragma unroll
    for (int u = 0; u < MAXP; ++u) {
        int P = wave + u * kWaves;
        if (P >= pairs) break;
        // slot -> (line t, column col); lanes of one column group take TL interleaved
        // lines and 16/TL consecutive columns
        int slot = P * 16 + c;
        bool valid = slot < slots;
        if (!valid) slot = 0;
        int b = slot / blk, w = slot - b * blk;
        int tl = w % TL, col = w / TL;
        int t = b * TL + tl;
        float2 x[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = lds[G.phys(t, col + (4 * g + i) * R16)];
        float yr[4], yi[4];
        dft16_mfma<PREC>(x, ops, lane, yr, yi);
        int q = col % s, p = col / s;
        ok[u] = valid;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int tq = 4 * g + i;
            float2 y = make_float2(yr[i] * m16, yi[i] * m16);
            if (p) y = cmul(y, a.stage_tw[p * tq * s]);
            res[u][i] = y;
            dst[u][i] = G.phys(t, q + s * (16 * p + tq));
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < MAXP; ++u) {
        if (!ok[u]) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) lds[dst[u][i]] = res[u][i];
    }
    __syncthreads();
}

// ------------------------------------------------------------------ VALU stages
// w16: cos(pi/8), sin(pi/8), cos(pi/4) (TileArgs::w16; radix 8 and 16 only)
template <int RADIX>
__device__ __forceinline__ void dft_small(float2 *v, int sign, const float *w16 = nullptr)
{
    if constexpr (RADIX == 2) {
        float2 a = v[0], b = v[1];
        v[0] = make_float2(a.x + b.x, a.y + b.y);
        v[1] = make_float2(a.x - b.x, a.y - b.y);
    } else if constexpr (RADIX == 4) {
        float2 s0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
        float2 d0 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
        float2 s1 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
        float2 d1 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
        // d1 * W_4 = d1 * (sign * i)
        float2 r = make_float2(-sign * d1.y, sign * d1.x);
        v[0] = make_float2(s0.x + s1.x, s0.y + s1.y);
        v[1] = make_float2(d0.x + r.x, d0.y + r.y);
        v[2] = make_float2(s0.x - s1.x, s0.y - s1.y);
        v[3] = make_float2(d0.x - r.x, d0.y - r.y);
    } else if constexpr (RADIX == 8) {
        // 8 = 2 x 4: x[j], j = 4*j1 + j2 ; X[k1 + 2*k2]
        const float h = w16[2];
        float2 a[4], b[4];
#pragma unroll
        for (int j2 = 0; j2 < 4; ++j2) {
            float2 x0 = v[j2], x1 = v[4 + j2];
            a[j2] = make_float2(x0.x + x1.x, x0.y + x1.y);  // k1 = 0
            b[j2] = make_float2(x0.x - x1.x, x0.y - x1.y);  // k1 = 1
        }
        // twiddle W_8^{j2} on the k1 = 1 branch
        b[1] = make_float2(h * (b[1].x - sign * b[1].y), h * (b[1].y + sign * b[1].x));
        b[2] = make_float2(-sign * b[2].y, sign * b[2].x);
        b[3] = make_float2(h * (-b[3].x - sign * b[3].y), h * (-b[3].y + sign * b[3].x));
        dft_small<4>(a, sign);
        dft_small<4>(b, sign);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            v[2 * k2] = a[k2];
            v[2 * k2 + 1] = b[k2];
        }
    } else if constexpr (RADIX == 16) {
        // 16 = 4 x 4: x[4*j1 + j2] ; X[k1 + 4*k2]
        float2 a[4][4];
#pragma unroll
        for (int j2 = 0; j2 < 4; ++j2) {
            float2 t[4] = {v[j2], v[4 + j2], v[8 + j2], v[12 + j2]};
            dft_small<4>(t, sign);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) a[k1][j2] = t[k1];
        }
        const float c1 = w16[0], s1 = w16[1];
        const float h = w16[2];
        // W_16^{j2*k1}: (k1,j2) in {1,2,3}^2
        const float wr[10] = {1.f, c1, h, s1, 0.f, -s1, -h, -c1, -1.f, -c1};
        const float wi[10] = {0.f, s1, h, c1, 1.f, c1, h, s1, 0.f, -s1};
#pragma unroll
        for (int k1 = 1; k1 < 4; ++k1)
#pragma unroll
            for (int j2 = 1; j2 < 4; ++j2) {
                int e = k1 * j2;
                float2 w = make_float2(wr[e], sign * wi[e]);
                a[k1][j2] = cmul(a[k1][j2], w);
            }
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            dft_small<4>(a[k1], sign);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = a[k1][k2];
        }
    }
}

// Odd radices (3, 5, 17, ...) as an O(r^2) DFT; W_r^e is read from the stage table
// (W_r^1 = W_R^(R/r)); the table index is uniform across the wave.
template <int RADIX>
__device__ __forceinline__ void dft_odd(float2 *v, const float2 *tw, int Rdiv)
{
    float2 o[RADIX];
#pragma unroll
    for (int t = 0; t < RADIX; ++t) {
        float2 acc = v[0];
#pragma unroll
        for (int j = 1; j < RADIX; ++j) {
            float2 w = tw[((j * t) % RADIX) * Rdiv];
            acc.x += v[j].x * w.x - v[j].y * w.y;
            acc.y += v[j].x * w.y + v[j].y * w.x;
        }
        o[t] = acc;
    }
#pragma unroll
    for (int t = 0; t < RADIX; ++t) v[t] = o[t];
}

// One radix-RADIX Stockham stage on the vector ALUs (fp32), in place in LDS.
template <int RADIX, class Geo>
__device__ void stage_valu(float2 *lds, const Geo &G, int s, float mul, const TileArgs &a)
{
    const int R = G.R, TL = G.TL;
    const int cols = R / RADIX;
    const int items = cols * G.T;
    constexpr int MAXI = (Geo::MAXP * kWaves * 256 + RADIX * kThreads - 1) / (RADIX * kThreads);
    float2 v[MAXI][RADIX];
#pragma unroll
    for (int k = 0; k < MAXI; ++k) {
        int I = threadIdx.x + k * kThreads;
        if (I >= items) break;
        int tl = I % TL, rest = I / TL;
        int col = rest % cols, b = rest / cols;
        int t = b * TL + tl;
#pragma unroll
        for (int j = 0; j < RADIX; ++j) v[k][j] = lds[G.phys(t, col + j * cols)];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MAXI; ++k) {
        int I = threadIdx.x + k * kThreads;
        if (I >= items) break;
        int tl = I % TL, rest = I / TL;
        int col = rest % cols, b = rest / cols;
        int t = b * TL + tl;
        int q = col % s, p = col / s;
        if constexpr (RADIX == 2 || RADIX == 4 || RADIX == 8 || RADIX == 16)
            dft_small<RADIX>(v[k], a.sign, a.w16);
        else
            dft_odd<RADIX>(v[k], a.stage_tw, cols);
#pragma unroll
        for (int tq = 0; tq < RADIX; ++tq) {
            float2 y = make_float2(v[k][tq].x * mul, v[k][tq].y * mul);
            if (p) y = cmul(y, a.stage_tw[p * tq * s]);
            lds[G.phys(t, q + s * (RADIX * p + tq))] = y;
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------------ copy in / out
struct TileOrigin {
    int l0b;
    int l1;
    int nvalid;  // lines of this tile that exist (the last tile along l0 may be partial)
    long long in_base, out_base;
};

__device__ __forceinline__ TileOrigin tile_origin(const TileArgs &a, long long tile)
{
    int tiles0 = (a.c0 + a.T - 1) / a.T;
    long long r = tile / tiles0;
    int t0 = (int)(tile - r * tiles0);
    int l1 = (int)(r % a.c1);
    long long l2 = r / a.c1;
    TileOrigin o;
    o.l0b = t0 * a.T;
    o.l1 = l1;
    o.nvalid = min(a.T, a.c0 - o.l0b);
    o.in_base = (long long)o.l0b * a.in_s0 + (long long)l1 * a.in_s1 + l2 * a.in_s2;
    o.out_base = (long long)o.l0b * a.out_s0 + (long long)l1 * a.out_s1 + l2 * a.out_s2;
    return o;
}

template <class Geo>
__device__ void copy_in(float2 *lds, const Geo &G, const TileArgs &a, const TileOrigin &o)
{
    const int R = G.R, T = G.T;
    const int *perm = a.in_perm;
    const int nv = o.nvalid;
    if (nv < T) {
        // zero the missing lines of a partial tile
        for (int e = threadIdx.x; e < (T - nv) * R; e += kThreads) {
            int t = nv + e / R, p = e % R;
            lds[G.phys(t, p)] = make_float2(0.f, 0.f);
        }
    }
    // Every thread first issues all of its global loads (UNR vectors), then scatters them
    // into LDS, so a tile's loads are in flight together.
    constexpr int UNR = 4;
    if (a.in_vec && a.in_ps == 1 && (R & 3) == 0 && (a.in_s0 & 3) == 0 && (o.in_base & 3) == 0) {
        // vectors of 4 consecutive points of one line
        const int per_line = R >> 2;
        const int nch = nv * per_line;
        for (int base = threadIdx.x; base < nch; base += UNR * kThreads) {
            float2 v[UNR][4];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                int ch = base + u * kThreads;
                if (ch < nch) {
                    int t = ch / per_line, p0 = (ch - t * per_line) << 2;
                    cload4(a.in, o.in_base + (long long)t * a.in_s0 + p0, a.in_fmt, v[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                int ch = base + u * kThreads;
                if (ch < nch) {
                    int t = ch / per_line, p0 = (ch - t * per_line) << 2;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        int pt = perm ? perm[p0 + k] : p0 + k;
                        lds[G.phys(t, pt)] = v[u][k];
                    }
                }
            }
        }
    } else if (a.in_vec && a.in_s0 == 1 && (nv & 3) == 0 && (o.in_base & 3) == 0 && (a.in_ps & 3) == 0) {
        // vectors of 4 consecutive lines at one point
        const int per_pt = nv >> 2;
        const int nch = R * per_pt;
        for (int base = threadIdx.x; base < nch; base += UNR * kThreads) {
            float2 v[UNR][4];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                int ch = base + u * kThreads;
                if (ch < nch) {
                    int p = ch / per_pt, t0 = (ch - p * per_pt) << 2;
                    cload4(a.in, o.in_base + (long long)p * a.in_ps + t0, a.in_fmt, v[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                int ch = base + u * kThreads;
                if (ch < nch) {
                    int p = ch / per_pt, t0 = (ch - p * per_pt) << 2;
                    int pt = perm ? perm[p] : p;
#pragma unroll
                    for (int k = 0; k < 4; ++k) lds[G.phys(t0 + k, pt)] = v[u][k];
                }
            }
        }
    } else {
        for (int e = threadIdx.x; e < nv * R; e += kThreads) {
            int t = e / R, p = e - t * R;
            float2 v = cload1(a.in, o.in_base + (long long)t * a.in_s0 + (long long)p * a.in_ps, a.in_fmt);
            lds[G.phys(t, perm ? perm[p] : p)] = v;
        }
    }
    __syncthreads();
}

template <class Geo>
__device__ __forceinline__ float2 out_value(const float2 *lds, const Geo &G, const TileArgs &a, int l0b, int l1,
                                            int t, int k)
{
    float2 v = lds[G.phys(t, k)];
    if (a.tw_N) {
        unsigned e = (unsigned)(a.tw_level ? l1 : l0b + t) * (unsigned)k;
        float2 w = cmul(a.tw_hi[e >> a.tw_bits], a.tw_lo[e & ((1u << a.tw_bits) - 1u)]);
        v = cmul(v, w);
    }
    v.x *= a.out_mul;
    v.y *= a.out_mul;
    return v;
}

template <class Geo>
__device__ void copy_out(const float2 *lds, const Geo &G, const TileArgs &a, const TileOrigin &o)
{
    const int R = G.R;
    const int nv = o.nvalid;
    constexpr int UNR = 4;
    if (a.out_vec && a.out_ps == 1 && (R & 3) == 0 && (a.out_s0 & 3) == 0 && (o.out_base & 3) == 0) {
        const int per_line = R >> 2;
        const int nch = nv * per_line;
        for (int base = threadIdx.x; base < nch; base += UNR * kThreads) {
            float2 v[UNR][4];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                int ch = base + u * kThreads;
                if (ch < nch) {
                    int t = ch / per_line, p0 = (ch - t * per_line) << 2;
#pragma unroll
                    for (int k = 0; k < 4; ++k) v[u][k] = out_value(lds, G, a, o.l0b, o.l1, t, p0 + k);
                }
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                int ch = base + u * kThreads;
                if (ch < nch) {
                    int t = ch / per_line, p0 = (ch - t * per_line) << 2;
                    cstore4(a.out, o.out_base + (long long)t * a.out_s0 + p0, a.out_fmt, v[u]);
                }
            }
        }
    } else if (a.out_vec && a.out_s0 == 1 && (nv & 3) == 0 && (o.out_base & 3) == 0 && (a.out_ps & 3) == 0) {
        const int per_pt = nv >> 2;
        const int nch = R * per_pt;
        for (int base = threadIdx.x; base < nch; base += UNR * kThreads) {
            float2 v[UNR][4];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                int ch = base + u * kThreads;
                if (ch < nch) {
                    int p = ch / per_pt, t0 = (ch - p * per_pt) << 2;
#pragma unroll
                    for (int k = 0; k < 4; ++k) v[u][k] = out_value(lds, G, a, o.l0b, o.l1, t0 + k, p);
                }
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                int ch = base + u * kThreads;
                if (ch < nch) {
                    int p = ch / per_pt, t0 = (ch - p * per_pt) << 2;
                    cstore4(a.out, o.out_base + (long long)p * a.out_ps + t0, a.out_fmt, v[u]);
                }
            }
        }
    } else {
        for (int e = threadIdx.x; e < nv * R; e += kThreads) {
            int t = e / R, p = e - t * R;
            cstore1(a.out, o.out_base + (long long)t * a.out_s0 + (long long)p * a.out_ps, a.out_fmt,
                    out_value(lds, G, a, o.l0b, o.l1, t, p));
        }
    }
}

// ------------------------------------------------------------------ the kernel
// Consecutive workgroups are dealt round-robin to the 8 XCDs; remap so that each XCD
// works on a contiguous run of tiles (neighbouring tiles share cache lines in strided
// passes, and each XCD has its own L2).
__device__ __forceinline__ long long xcd_remap(long long bid, long long nblk)
{
    if ((nblk & 7) != 0) return bid;
    return (bid & 7) * (nblk >> 3) + (bid >> 3);
}

// Power-of-two line length R = 16^k * tail (tail in {1, 2, 4, 8}) with compile-time
// geometry: tail stage first (VALU, fp32), then k radix-16 matrix-core stages.
template <int PREC, int R, int TL, int T>
__global__ void __launch_bounds__(kThreads) fft_tile_pow2(TileArgs a)
{
    extern __shared__ float2 lds[];
    GeoC<R, T, TL> G;
    constexpr int L2 = __builtin_ctz(R);
    constexpr int NS16 = L2 / 4;
    constexpr int TAIL = 1 << (L2 % 4);
    long long tile = xcd_remap(blockIdx.x, gridDim.x);
    TileOrigin o = tile_origin(a, tile);
    copy_in(lds, G, a, o);
    int st = 0, s = 1;
    if constexpr (TAIL > 1) {
        stage_valu<TAIL>(lds, G, 1, a.stage_mul[0], a);
        st = 1;
        s = TAIL;
    }
#pragma unroll
    for (int k = 0; k < NS16; ++k) {
        if constexpr (PREC == PREC_FP32)
            stage_valu<16>(lds, G, s, a.stage_mul[st], a);
        else
            stage16_mfma<PREC>(lds, G, s, a.stage_mul[st], a);
        ++st;
        s *= 16;
    }
    copy_out(lds, G, a, o);
}

// Any line length made of radices {16, 2, 3, 4, 5, 8, 17} (runtime schedule): the body
// of one tile, shared by the per-plan kernel and the work-queue kernel.
template <int PREC>
__device__ __forceinline__ void generic_tile(const TileArgs &a, long long tile, float2 *lds)
{
    GeoR G{a.R, a.T, a.TL};
    TileOrigin o = tile_origin(a, tile);
    copy_in(lds, G, a, o);
    int s = 1;
    for (int st = 0; st < a.nstages; ++st) {
        int r = a.radix[st];
        float mul = a.stage_mul[st];
        switch (r) {
        case 16:
            if constexpr (PREC == PREC_FP32)
                stage_valu<16>(lds, G, s, mul, a);
            else
                stage16_mfma<PREC>(lds, G, s, mul, a);
            break;
        case 2: stage_valu<2>(lds, G, s, mul, a); break;
        case 3: stage_valu<3>(lds, G, s, mul, a); break;
        case 4: stage_valu<4>(lds, G, s, mul, a); break;
        case 5: stage_valu<5>(lds, G, s, mul, a); break;
        case 8: stage_valu<8>(lds, G, s, mul, a); break;
        default: stage_valu<17>(lds, G, s, mul, a); break;
        }
        s *= r;
    }
    copy_out(lds, G, a, o);
}

template <int PREC>
__global__ void __launch_bounds__(kThreads) fft_tile_generic(TileArgs a)
{
    extern __shared__ float2 lds[];
    generic_tile<PREC>(a, xcd_remap(blockIdx.x, gridDim.x), lds);
}

// Many transforms of uneven length and batch count in one launch: jobs[j] describes one
// request (its tiles are first_tile[j] .. first_tile[j+1]-1); every workgroup takes the
// next tile from a device-side counter until the queue is empty, so each wave ends once
// all tiles are handed out.
template <int PREC>
__global__ void __launch_bounds__(kThreads) fft_queue_kernel(const TileArgs *jobs, const long long *first_tile,
                                                             int njobs, long long total, unsigned long long *counter)
{
    extern __shared__ float2 lds[];
    __shared__ long long s_tile;
    __shared__ int s_job;
    for (;;) {
        if (threadIdx.x == 0) {
            long long t = (long long)atomicAdd(counter, 1ull);
            int lo = 0, hi = njobs - 1;
            while (lo < hi) {  // last job with first_tile <= t
                int mid = (lo + hi + 1) >> 1;
                if (first_tile[mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            s_tile = t;
            s_job = lo;
        }
        __syncthreads();
        const long long t = s_tile;
        const int j = s_job;
        __syncthreads();  // every thread has read the shared slot before it is reused
        if (t >= total) break;
        generic_tile<PREC>(jobs[j], t - first_tile[j], lds);
        __syncthreads();
    }
}

__global__ void queue_reset_kernel(unsigned long long *counter) { *counter = 0ull; }

hipError_t launch_queue(int prec, const TileArgs *jobs, const long long *first_tile, int njobs, long long total,
                        unsigned long long *counter, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(1), 0, s, counter);
    size_t lds = (size_t)tile_lds_bytes(kMaxTile);
    switch (prec) {
    case PREC_FP16:
        hipLaunchKernelGGL(fft_queue_kernel<PREC_FP16>, dim3(grid), dim3(kThreads), lds, s, jobs, first_tile, njobs,
                           total, counter);
        break;
    case PREC_BF16:
        hipLaunchKernelGGL(fft_queue_kernel<PREC_BF16>, dim3(grid), dim3(kThreads), lds, s, jobs, first_tile, njobs,
                           total, counter);
        break;
    case PREC_SPLIT:
        hipLaunchKernelGGL(fft_queue_kernel<PREC_SPLIT>, dim3(grid), dim3(kThreads), lds, s, jobs, first_tile, njobs,
                           total, counter);
        break;
    default:
        hipLaunchKernelGGL(fft_queue_kernel<PREC_FP32>, dim3(grid), dim3(kThreads), lds, s, jobs, first_tile, njobs,
                           total, counter);
        break;
    }
    return hipGetLastError();
}

int queue_grid(int prec)
{
    int dev = 0, ncu = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 256;
    const void *fn = prec == PREC_FP16    ? (const void *)fft_queue_kernel<PREC_FP16>
                     : prec == PREC_BF16  ? (const void *)fft_queue_kernel<PREC_BF16>
                     : prec == PREC_SPLIT ? (const void *)fft_queue_kernel<PREC_SPLIT>
                                          : (const void *)fft_queue_kernel<PREC_FP32>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kThreads, tile_lds_bytes(kMaxTile)) != hipSuccess || nb < 1)
        nb = 1;
    return nb * ncu;
}

// ------------------------------------------------------------------ launch table
typedef void (*tile_fn)(TileArgs);

template <int PREC, int R, int TL>
constexpr tile_fn pow2_fn()
{
    return fft_tile_pow2<PREC, R, TL, (kMaxTile / R > TL ? kMaxTile / R : TL)>;
}

template <int PREC>
tile_fn pick_pow2(int R, int TL)
{
    if (TL == 16) {
        switch (R) {
        case 16: return pow2_fn<PREC, 16, 16>();
        case 32: return pow2_fn<PREC, 32, 16>();
        case 64: return pow2_fn<PREC, 64, 16>();
        case 128: return pow2_fn<PREC, 128, 16>();
        case 256: return pow2_fn<PREC, 256, 16>();
        default: return nullptr;
        }
    }
    if (TL == 1) {
        switch (R) {
        case 16: return pow2_fn<PREC, 16, 1>();
        case 32: return pow2_fn<PREC, 32, 1>();
        case 64: return pow2_fn<PREC, 64, 1>();
        case 128: return pow2_fn<PREC, 128, 1>();
        case 256: return pow2_fn<PREC, 256, 1>();
        case 512: return pow2_fn<PREC, 512, 1>();
        case 1024: return pow2_fn<PREC, 1024, 1>();
        case 2048: return pow2_fn<PREC, 2048, 1>();
        case 4096: return pow2_fn<PREC, 4096, 1>();
        default: return nullptr;
        }
    }
    return nullptr;
}

// Tile size (lines per workgroup) of the compiled power-of-two kernels.
int pow2_tile_lines(int R, int TL)
{
    int t = kMaxTile / R;
    return t > TL ? t : TL;
}

tile_fn find_tile_kernel(int prec, int R, int TL, bool generic)
{
    if (!generic) {
        switch (prec) {
        case PREC_FP16: return pick_pow2<PREC_FP16>(R, TL);
        case PREC_BF16: return pick_pow2<PREC_BF16>(R, TL);
        case PREC_SPLIT: return pick_pow2<PREC_SPLIT>(R, TL);
        case PREC_FP32: return pick_pow2<PREC_FP32>(R, TL);
        default: return nullptr;
        }
    }
    switch (prec) {
    case PREC_FP16: return fft_tile_generic<PREC_FP16>;
    case PREC_BF16: return fft_tile_generic<PREC_BF16>;
    case PREC_SPLIT: return fft_tile_generic<PREC_SPLIT>;
    case PREC_FP32: return fft_tile_generic<PREC_FP32>;
    default: return nullptr;
    }
}

hipError_t launch_tile(const void *fn, const TileArgs &a, long long ntiles, hipStream_t stream, hipEvent_t start,
                       hipEvent_t stop)
{
    size_t lds = (size_t)tile_lds_bytes(a.R * a.T);
    dim3 grid((unsigned)ntiles), block(kThreads);
    void *args[] = {const_cast<TileArgs *>(&a)};
    if (start || stop) return hipExtLaunchKernel(fn, grid, block, args, lds, stream, start, stop, 0);
    return hipLaunchKernel(fn, grid, block, args, lds, stream);
}

hipError_t prepare_tile_kernel(const void *fn, int R, int T)
{
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, tile_lds_bytes(R * T));
}

}  // namespace mfft
