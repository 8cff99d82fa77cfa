// Streaming kernel for short strided DFTs -- the small first pass of a four-step split
// along a strided axis: R = 2, 4, 8, 16 (e.g. 1024 = 4 x 256 columns of a 2D transform)
// and the odd R = 3, 5, 9, 15, 17 that take the odd factor of lengths odd * 2^a (e.g.
// 2176 = 17 x 128), so the power-of-two rest runs on the matrix-core kernels. One thread
// owns one line: it loads the R points (consecutive threads take consecutive lines, so
// every load is coalesced), transforms them in fp32 registers, applies the four-step
// twiddle and the stage scale, and stores -- no LDS, one HBM pass. A thread reads and
// writes the same R positions, so the pass may run in place.
#include "mfft_common.h"

namespace mfft {
namespace small {

template <int FMT>
constexpr int elem_bytes() { return FMT == FMT_C32 ? 8 : 4; }

template <int FMT>
__device__ __forceinline__ float2 unpack(uint32_t w)
{
    if constexpr (FMT == FMT_C16)
        return make_float2((float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffff)),
                           (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)));
    else
        return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

template <int FMT>
__device__ __forceinline__ float2 ld(const void *p, long long i)
{
    if constexpr (FMT == FMT_C32)
        return reinterpret_cast<const float2 *>(p)[i];
    else
        return unpack<FMT>(reinterpret_cast<const uint32_t *>(p)[i]);
}

typedef unsigned int u2v __attribute__((ext_vector_type(2)));

// A buffer resource over the first `bytes` bytes of p (a launch whose offsets are all below
// 2^32 bytes): an element's address is the resource base plus a per-lane 32-bit offset plus
// a uniform offset, with no 64-bit arithmetic per element.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t span_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}

template <int FMT>
__device__ __forceinline__ float2 ld_off(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned uoff)
{
    if constexpr (FMT == FMT_C32) {
        const u2v w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)uoff, 0);
        return make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
    } else {
        return unpack<FMT>(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)uoff, 0));
    }
}

template <int FMT>
__device__ __forceinline__ uint32_t pack(float2 v)
{
    if constexpr (FMT == FMT_C16)
        return pack_c16(v.x, v.y);
    else
        return pack_cbf16(v.x, v.y);
}

template <int FMT>
__device__ __forceinline__ void st(void *p, long long i, float2 v)
{
    if constexpr (FMT == FMT_C32)
        reinterpret_cast<float2 *>(p)[i] = v;
    else
        reinterpret_cast<uint32_t *>(p)[i] = pack<FMT>(v);
}

template <int FMT>
__device__ __forceinline__ void st_off(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned uoff, float2 v)
{
    if constexpr (FMT == FMT_C32) {
        u2v w;
        w.x = __float_as_uint(v.x);
        w.y = __float_as_uint(v.y);
        __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)off, (int)uoff, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(pack<FMT>(v), r, (int)off, (int)uoff, 0);
    }
}

__device__ __forceinline__ float2 cm(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ void dft4(float2 *v, int sg)
{
    float2 s0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y), d0 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    float2 s1 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y), d1 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    float2 r = make_float2(-sg * d1.y, sg * d1.x);
    v[0] = make_float2(s0.x + s1.x, s0.y + s1.y);
    v[1] = make_float2(d0.x + r.x, d0.y + r.y);
    v[2] = make_float2(s0.x - s1.x, s0.y - s1.y);
    v[3] = make_float2(d0.x - r.x, d0.y - r.y);
}

// DFT of R points in place, fp32; w16 = cos(pi/8), sin(pi/8), cos(pi/4) (TileArgs::w16),
// used by R = 8 and 16.
template <int R>
__device__ __forceinline__ void dft(float2 *v, int sg, const float *w16)
{
    if constexpr (R == 2) {
        float2 a = v[0], b = v[1];
        v[0] = make_float2(a.x + b.x, a.y + b.y);
        v[1] = make_float2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        dft4(v, sg);
    } else {
        // R = 2 * H (H = 4 or 8): radix-2 butterfly, W_R^j on the odd half, two DFT_H
        constexpr int H = R / 2;
        float2 a[H], b[H];
#pragma unroll
        for (int j = 0; j < H; ++j) {
            a[j] = make_float2(v[j].x + v[j + H].x, v[j].y + v[j + H].y);
            b[j] = make_float2(v[j].x - v[j + H].x, v[j].y - v[j + H].y);
        }
        const float c1 = w16[0], s1 = w16[1], h = w16[2];
        // W_16^e (cos, sin) for e = 0..7; W_8^e = W_16^(2e)
        const float wc[8] = {1.f, c1, h, s1, 0.f, -s1, -h, -c1};
        const float ws[8] = {0.f, s1, h, c1, 1.f, c1, h, s1};
#pragma unroll
        for (int j = 1; j < H; ++j) {
            const int e = j * (16 / R);
            b[j] = cm(b[j], make_float2(wc[e], sg * ws[e]));
        }
        dft<H>(a, sg, w16);
        dft<H>(b, sg, w16);
#pragma unroll
        for (int k = 0; k < H; ++k) {
            v[2 * k] = a[k];
            v[2 * k + 1] = b[k];
        }
    }
}

// Odd R: y_k = x_0 + sum_j S_j c_jk + i sum_j D_j t_jk and y_{R-k} = x_0 + sum_j S_j c_jk -
// i sum_j D_j t_jk (j, k = 1..(R-1)/2), with S_j = x_j + x_{R-j}, D_j = x_j - x_{R-j} and
// W_R^m = c_m + i t_m from the pass table (FP64-generated, carries the transform's sign):
// (R-1)^2 real multiply-adds per line instead of 4 (R-1)^2.
template <int R>
__device__ __forceinline__ void dft_odd(float2 *v, const float2 *__restrict__ tab)
{
    constexpr int H = (R - 1) / 2;
    float2 w[R];
#pragma unroll
    for (int m = 0; m < R; ++m) w[m] = tab[m];
    float2 S[H + 1], D[H + 1];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        S[j] = make_float2(v[j].x + v[R - j].x, v[j].y + v[R - j].y);
        D[j] = make_float2(v[j].x - v[R - j].x, v[j].y - v[R - j].y);
    }
    float2 y0 = v[0];
#pragma unroll
    for (int j = 1; j <= H; ++j) y0 = make_float2(y0.x + S[j].x, y0.y + S[j].y);
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        float2 a = v[0], b = make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const float2 t = w[(j * k) % R];
            a.x += S[j].x * t.x;
            a.y += S[j].y * t.x;
            b.x += D[j].x * t.y;
            b.y += D[j].y * t.y;
        }
        // i * b = (-b.y, b.x)
        v[k] = make_float2(a.x - b.y, a.y + b.x);
        v[R - k] = make_float2(a.x + b.y, a.y - b.x);
    }
    v[0] = y0;
}

// Stage scale, four-step twiddle W_N^(mult k) and store of one line. W_N^mult comes from
// the two-level table (FP64-generated) and its powers by repeated products (fewer than
// R <= 17 fp32 roundings; two table reads per line instead of two per point); each point is
// stored as soon as it is ready. OFF32 (every output offset of the launch below 2^32 bytes,
// out_span): stores through a buffer resource, and the stage scale folded into the twiddle
// powers (one complex product per point).
template <int R, int OUTF, bool OFF32>
__device__ __forceinline__ void finish(float2 *v, const TileArgs &a, int l0, int l1, long long ob)
{
    const unsigned mult = (unsigned)(a.tw_level ? l1 : l0);
    if constexpr (OFF32) {
        constexpr unsigned EB = elem_bytes<OUTF>();
        const __amdgpu_buffer_rsrc_t r = span_rsrc(a.out, a.out_span);
        const unsigned off = (unsigned)ob * EB, step = (unsigned)a.out_ps * EB;
        st_off<OUTF>(r, off, 0u, make_float2(v[0].x * a.last_mul, v[0].y * a.last_mul));
        if (a.tw_N) {
            const float2 w1 = cm(a.tw_hi[mult >> a.tw_bits], a.tw_lo[mult & ((1u << a.tw_bits) - 1u)]);
            float2 u = make_float2(w1.x * a.last_mul, w1.y * a.last_mul);  // last_mul W_N^(mult k)
#pragma unroll
            for (int k = 1; k < R; ++k) {
                if (k > 1) u = cm(u, w1);
                st_off<OUTF>(r, off, k * step, cm(v[k], u));
            }
        } else {
#pragma unroll
            for (int k = 1; k < R; ++k)
                st_off<OUTF>(r, off, k * step, make_float2(v[k].x * a.last_mul, v[k].y * a.last_mul));
        }
    } else {
        float2 w1 = make_float2(1.f, 0.f), wk = w1;
        if (a.tw_N) w1 = cm(a.tw_hi[mult >> a.tw_bits], a.tw_lo[mult & ((1u << a.tw_bits) - 1u)]);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            float2 y = make_float2(v[k].x * a.last_mul, v[k].y * a.last_mul);
            if (a.tw_N && k) {
                wk = k == 1 ? w1 : cm(wk, w1);
                y = cm(y, wk);
            }
            st<OUTF>(a.out, ob + (long long)k * a.out_ps, y);
        }
    }
}

// Lines are (l0, l1, l2) as in TileArgs; one thread per line.
template <int R, int INF, int OUTF>
__global__ void __launch_bounds__(256) small_cols(TileArgs a)
{
    const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long nlines = (long long)a.c0 * a.c1;  // lines per transform
    const long long nl = nlines * a.c2;
    // line index -> (l0, l1, l2); 32-bit divisions whenever the line count allows
    int l0, l1;
    long long l2;
    if (a.in_mul && in_mul_threads(nlines, a.c2) <= 0xffffffffLL) {
        // conv filter product: 64 consecutive lines of one transform per wave, the same lines
        // of the launch's other transforms (l2) in the waves that follow (the launch covers
        // the lines in whole groups of 64). A wave then stores 64 consecutive words, and the
        // filter-spectrum words of its lines, read by the c2 waves of a group close in time,
        // come from memory once per launch.
        if (id >= in_mul_threads(nlines, a.c2)) return;
        const unsigned u = (unsigned)id, c2 = (unsigned)a.c2, c0 = (unsigned)a.c0;
        const unsigned grp = u / (64u * c2), rem = u - grp * (64u * c2);
        const unsigned line = grp * 64u + (rem & 63u);
        if (line >= (unsigned)nlines) return;
        l2 = rem >> 6;
        const unsigned q = line / c0;
        l0 = (int)(line - q * c0);
        l1 = (int)q;
    } else if (id >= nl) {
        return;
    } else if (nl <= 0xffffffffLL) {
        const unsigned u = (unsigned)id, c0 = (unsigned)a.c0, c1 = (unsigned)a.c1;
        const unsigned r = u / c0;
        l0 = (int)(u - r * c0);
        const unsigned q = r / c1;
        l1 = (int)(r - q * c1);
        l2 = q;
    } else {
        const long long r = id / a.c0;
        l0 = (int)(id - r * a.c0);
        l1 = (int)(r % a.c1);
        l2 = r / a.c1;
    }
    const long long ib = (long long)l0 * a.in_s0 + (long long)l1 * a.in_s1 + l2 * a.in_s2;
    const long long ob = (long long)l0 * a.out_s0 + (long long)l1 * a.out_s1 + l2 * a.out_s2;
    float2 v[R];
    if (a.real_src) {
        // conv: two real images of a pair, zero-padded (replaces a separate packing pass).
        // A line lies in one padded transform: its first point is decoded once into
        // (transform, y, x) and the others follow by the point stride split into rows and
        // columns (dx < padded width, so x carries at most once per step).
        const long long tr = ib / a.real_p;
        const long long rem = ib - tr * a.real_p;
        int y = (int)(rem / a.real_pw), x = (int)(rem - (long long)y * a.real_pw);
        const int dy = (int)(a.in_ps / a.real_pw), dx = (int)(a.in_ps - (long long)dy * a.real_pw);
        const long long t = a.batch_base + tr;
        const long long k = t / a.real_c, c = t - k * a.real_c;
        const long long hw = (long long)a.real_h * a.real_w;
        const long long i0 = (2 * k * a.real_c + c) * hw;
        const bool second = a.real_src2 || 2 * k + 1 < a.real_b;
        const void *src1 = a.real_src2 ? a.real_src2 : a.real_src;
        const long long i1 = a.real_src2 ? i0 : ((2 * k + 1) * a.real_c + c) * hw;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            float2 val = make_float2(0.f, 0.f);
            if (y < a.real_h && x < a.real_w) {
                const long long off = (long long)y * a.real_w + x;
                if (a.real_f64) {
                    val.x = (float)reinterpret_cast<const double *>(a.real_src)[i0 + off];
                    if (second) val.y = (float)reinterpret_cast<const double *>(src1)[i1 + off];
                } else {
                    val.x = reinterpret_cast<const float *>(a.real_src)[i0 + off];
                    if (second) val.y = reinterpret_cast<const float *>(src1)[i1 + off];
                }
            }
            v[j] = val;
            x += dx;
            y += dy;
            if (x >= a.real_pw) {
                x -= a.real_pw;
                ++y;
            }
        }
    } else if ((a.off32 & 1) && !a.in_mul) {
        // every input offset of the launch is below 2^32 bytes
        constexpr unsigned EB = elem_bytes<INF>();
        const __amdgpu_buffer_rsrc_t r = span_rsrc(a.in, a.in_span);
        const unsigned off = (unsigned)ib * EB, step = (unsigned)a.in_ps * EB;
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = ld_off<INF>(r, off, j * step);
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = ld<INF>(a.in, ib + (long long)j * a.in_ps);
    }
    if (a.in_mul) {
        // conv: the pointwise product with the resident filter spectrum, fused into the load
        // (the line's points stay in one period, so the index is reduced once)
        long long m = ib % a.in_mul_period;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            v[j] = cm(v[j], a.in_mul[m]);
            m += a.in_ps;
            if (m >= a.in_mul_period) m -= a.in_mul_period;
        }
    }
    if constexpr (R & 1)
        dft_odd<R>(v, a.stage_tw);
    else
        dft<R>(v, a.sign, a.w16);
    // the fused conv passes (image load, filter product) keep 64-bit addresses: with the
    // buffer form they ran 2-4% slower, the plain passes 14% faster
    if ((a.off32 & 2) && !a.in_mul && !a.real_src)
        finish<R, OUTF, true>(v, a, l0, l1, ob);
    else
        finish<R, OUTF, false>(v, a, l0, l1, ob);
}

template <int INF, int OUTF>
const void *pick(int R)
{
    switch (R) {
    case 2: return (const void *)small_cols<2, INF, OUTF>;
    case 4: return (const void *)small_cols<4, INF, OUTF>;
    case 8: return (const void *)small_cols<8, INF, OUTF>;
    case 16: return (const void *)small_cols<16, INF, OUTF>;
    case 3: return (const void *)small_cols<3, INF, OUTF>;
    case 5: return (const void *)small_cols<5, INF, OUTF>;
    case 9: return (const void *)small_cols<9, INF, OUTF>;
    case 15: return (const void *)small_cols<15, INF, OUTF>;
    case 17: return (const void *)small_cols<17, INF, OUTF>;
    default: return nullptr;
    }
}

}  // namespace small

const void *find_small_kernel(int R, int inf, int outf)
{
    using namespace small;
    if (inf == FMT_C16 && outf == FMT_C16) return pick<FMT_C16, FMT_C16>(R);
    if (inf == FMT_C16 && outf == FMT_C32) return pick<FMT_C16, FMT_C32>(R);
    if (inf == FMT_C32 && outf == FMT_C32) return pick<FMT_C32, FMT_C32>(R);
    if (inf == FMT_C32 && outf == FMT_C16) return pick<FMT_C32, FMT_C16>(R);
    if (inf == FMT_CBF16 && outf == FMT_CBF16) return pick<FMT_CBF16, FMT_CBF16>(R);
    return nullptr;
}

}  // namespace mfft
