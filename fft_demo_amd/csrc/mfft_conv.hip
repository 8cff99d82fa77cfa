// Elementwise kernels of the FFT convolution operator (pack/pad, pointwise multiply-
// accumulate, crop/unpack). The transforms themselves are mfft 2D plans.
//
// Reference steps these replace (tensorcore_fft_conv2d_kernel.cu): filter rotation
// (fft_conv2d_inverse_filter :1234-1253), real->complex padding with digit reversal
// (fft_conv2d_bit_rev_real_pad_to_cpx_double2half :1222-1232), the channel sum
// (fft_cpx_point_wise_mul_add :1033-1056), the per-filter power-of-10 quantisation
// (quant_p1/p2/p3 :1058-1098; only the reference-method mode uses it -- otherwise spectra
// stay fp32 or the transforms are scaled per stage), the dequantisation
// (fft_cpx_point_wise_dequant :1129-1139) and the crop/scale readout
// (read_out_scale_double2double_kernel :1390-1397).
#include "mfft_common.h"

namespace mfft {

// Image and product spectra are complex fp32 (FMT_C32) or fp16 (FMT_C16) per operator.
template <int FMT>
__device__ __forceinline__ float2 spec_ld(const void *p, long long i)
{
    if constexpr (FMT == FMT_C32) {
        return reinterpret_cast<const float2 *>(p)[i];
    } else {
        const uint32_t w = reinterpret_cast<const uint32_t *>(p)[i];
        return make_float2((float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)),
                           (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)));
    }
}

template <int FMT>
__device__ __forceinline__ void spec_st(void *p, long long i, float2 v)
{
    if constexpr (FMT == FMT_C32)
        reinterpret_cast<float2 *>(p)[i] = v;
    else
        reinterpret_cast<uint32_t *>(p)[i] = pack_c16(v.x, v.y);
}

// Two real images (batch items 2k and 2k+1, same channel) become the real and imaginary
// parts of one complex image, zero-padded to Ph x Pw at the origin.
// src: real [B][C][h][w] (fp32 or fp64); dst: complex [pairs][C][Ph][Pw] (FMT).
template <typename T, int FMT>
__global__ void pack_pairs_kernel(const T *src, void *dst, int B, int C, int h, int w, int Ph, int Pw, long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int x = (int)(i % Pw);
    long long r = i / Pw;
    int y = (int)(r % Ph);
    r /= Ph;
    int c = (int)(r % C);
    long long k = r / C;
    float2 v = make_float2(0.f, 0.f);
    if (y < h && x < w) {
        long long b0 = 2 * k, b1 = 2 * k + 1;
        long long off = ((long long)c * h + y) * w + x;
        v.x = (float)src[b0 * C * h * w + off];
        if (b1 < B) v.y = (float)src[b1 * C * h * w + off];
    }
    spec_st<FMT>(dst, i, v);
}

// Filter [F][C][fh][fw] real -> complex fp32 [F][C][Ph][Pw], optionally rotated by 180
// degrees (convolution by the rotated filter = cross-correlation).
template <typename T>
__global__ void pack_filters_kernel(const T *src, float2 *dst, int fh, int fw, int Ph, int Pw, int rotate, long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int x = (int)(i % Pw);
    long long r = i / Pw;
    int y = (int)(r % Ph);
    long long fc = r / Ph;
    float2 v = make_float2(0.f, 0.f);
    if (y < fh && x < fw) {
        int sy = rotate ? fh - 1 - y : y, sx = rotate ? fw - 1 - x : x;
        v.x = (float)src[(fc * fh + sy) * fw + sx];
    }
    dst[i] = v;
}

// Y[k][f][p] = sum_c X[k][c][p] * H[f][c][p]   (X, Y: FMT; H: fp32)
template <int XFMT, int YFMT>
__global__ void pointwise_mac_kernel(const void *X, const float2 *H, void *Y, int C, int F, long long P,
                                     long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    long long p = i % P;
    long long r = i / P;
    int f = (int)(r % F);
    long long k = r / F;
    float re = 0.f, im = 0.f;
    for (int c = 0; c < C; ++c) {
        float2 a = spec_ld<XFMT>(X, (k * C + c) * P + p);
        float2 b = H[((long long)f * C + c) * P + p];
        re += a.x * b.x - a.y * b.y;
        im += a.x * b.y + a.y * b.x;
    }
    spec_st<YFMT>(Y, i, make_float2(re, im));
}

// The reference method's quantisation of the product spectra (quant_p1/p2/p3,
// tensorcore_fft_conv2d_kernel.cu:1058-1098): per output map (pair, filter) the largest
// |re| or |im| m over its P values sets the scale s = 1 if m <= 1000, else 10^(trunc(log10 m)
// - 1), and the map is stored as fp16(R / s) for the fp16 inverse transform. Step 1: each
// workgroup takes a chunk of one map and writes the chunk's maximum.
__global__ void quant_max_kernel(const float2 *R, float *partial, long long P, int nchunk)
{
    __shared__ float red[256];
    const long long map = blockIdx.x / nchunk;
    const int chunk = (int)(blockIdx.x - map * nchunk);
    const long long per = (P + nchunk - 1) / nchunk;
    const long long b = chunk * per, e = b + per < P ? b + per : P;
    float m = 0.f;
    for (long long i = b + threadIdx.x; i < e; i += 256) {
        const float2 v = R[map * P + i];
        m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// The quantisation scale of a map from its largest magnitude (quant_p2's rule; the power of
// ten by repeated products in fp32, exact up to 10^10). A map whose forward overflowed (an
// infinite or NaN maximum) keeps scale 1, so the overflow shows in its output as it does in
// the reference's; the loop is bounded by the largest finite maximum (k <= 37).
__device__ __forceinline__ float quant_scale(float m)
{
    if (!(m > 1000.f) || !(m <= 3.4e38f)) return 1.f;
    const int k = (int)log10f(m) - 1;
    float s = 1.f;
    for (int j = 0; j < k; ++j) s *= 10.f;
    return s;
}

// Step 2: each workgroup reduces its map's chunk maxima, takes the scale (chunk 0 records
// it for the dequantisation) and stores its chunk as fp16(R / s).
__global__ void quant_apply_kernel(const float2 *R, uint32_t *Q, const float *partial, float *scale, long long P,
                                   int nchunk)
{
    const long long map = blockIdx.x / nchunk;
    const int chunk = (int)(blockIdx.x - map * nchunk);
    float m = 0.f;
    for (int c = 0; c < nchunk; ++c) m = fmaxf(m, partial[map * nchunk + c]);
    const float s = quant_scale(m), inv = 1.f / s;
    if (chunk == 0 && threadIdx.x == 0) scale[map] = s;
    const long long per = (P + nchunk - 1) / nchunk;
    const long long b = chunk * per, e = b + per < P ? b + per : P;
    for (long long i = b + threadIdx.x; i < e; i += 256) {
        const float2 v = R[map * P + i];
        Q[map * P + i] = pack_c16(v.x * inv, v.y * inv);
    }
}

// out[2k][f][y][x] = scale * Re Y[k][f][y + oy][x + ox], out[2k+1][f][y][x] = scale * Im ...
// (each spectrum value is read once for both images). Workgroup b covers 256 * V columns of
// output row b / xb (xb workgroups per row), so the row is decoded once per workgroup; a
// thread takes V consecutive columns (V = 4: one 16-byte store per image, for fp32 output
// rows whose width is a multiple of 4 on a 16-byte aligned buffer).
template <typename T, int FMT, int V>
__global__ void crop_unpack_kernel(const void *Y, T *out, int B, int F, int oh, int ow, int Ph, int Pw, int oy,
                                   int ox, float scale, int xb, const float *map_scale)
{
    const unsigned row = blockIdx.x / (unsigned)xb;
    const int x = ((int)(blockIdx.x - row * (unsigned)xb) * 256 + threadIdx.x) * V;
    if (x >= ow) return;
    const unsigned kf = row / (unsigned)oh;
    const int y = (int)(row - kf * (unsigned)oh);
    const long long k = kf / (unsigned)F;
    const int f = (int)(kf - (unsigned)k * (unsigned)F);
    const long long src = ((k * F + f) * Ph + (y + oy)) * (long long)Pw + (x + ox);
    const long long plane = (long long)F * oh * ow, off = ((long long)f * oh + y) * ow + x;
    const bool second = 2 * k + 1 < B;
    if (map_scale) scale *= map_scale[k * F + f];  // the reference method's dequantisation
    if constexpr (V == 1) {
        const float2 v = spec_ld<FMT>(Y, src);
        out[2 * k * plane + off] = (T)(v.x * scale);
        if (second) out[(2 * k + 1) * plane + off] = (T)(v.y * scale);
    } else {
        static_assert(V == 4, "4 columns per thread");
        float re[4], im[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float2 v = spec_ld<FMT>(Y, src + j);
            re[j] = v.x * scale;
            im[j] = v.y * scale;
        }
        *reinterpret_cast<float4 *>(out + 2 * k * plane + off) = make_float4(re[0], re[1], re[2], re[3]);
        if (second)
            *reinterpret_cast<float4 *>(out + (2 * k + 1) * plane + off) = make_float4(im[0], im[1], im[2], im[3]);
    }
}

// Two real images a, b (separate buffers, [h][w]) -> one complex image a + i b, zero-padded.
template <typename T, int FMT>
__global__ void pack_two_kernel(const T *a, const T *b, void *dst, int h, int w, int Ph, int Pw, long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int x = (int)(i % Pw), y = (int)(i / Pw);
    float2 v = make_float2(0.f, 0.f);
    if (y < h && x < w) {
        long long off = (long long)y * w + x;
        v = make_float2((float)a[off], (float)b[off]);
    }
    spec_st<FMT>(dst, i, v);
}

// Spectrum X of a + i b (a, b real): A = (X(k) + conj X(-k)) / 2 and B = (X(k) - conj X(-k))
// / 2i are the spectra of a and b, so Y = A H0 + i B H1 = ((X + X*)/2) H0 + ((X - X*)/2) H1
// (X* = conj X(-k)) transforms back to (a conv h0) + i (b conv h1): two convolutions with
// different filters from one forward and one inverse transform.
template <int FMT>
__global__ void pointwise_two_kernel(const void *X, const float2 *H0, const float2 *H1, void *Y, int Ph, int Pw)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)Ph * Pw) return;
    int v = (int)(i % Pw), u = (int)(i / Pw);
    int uq = u ? Ph - u : 0, vq = v ? Pw - v : 0;
    float2 xp = spec_ld<FMT>(X, i), xq = spec_ld<FMT>(X, (long long)uq * Pw + vq);
    xq.y = -xq.y;
    float2 s = make_float2(0.5f * (xp.x + xq.x), 0.5f * (xp.y + xq.y));
    float2 d = make_float2(0.5f * (xp.x - xq.x), 0.5f * (xp.y - xq.y));
    float2 h0 = H0[i], h1 = H1[i];
    spec_st<FMT>(Y, i,
                 make_float2(s.x * h0.x - s.y * h0.y + d.x * h1.x - d.y * h1.y,
                             s.x * h0.y + s.y * h0.x + d.x * h1.y + d.y * h1.x));
}

// out0 = Re Y, out1 = Im Y over the output window at (oy, ox).
template <typename T, int FMT>
__global__ void crop_two_kernel(const void *Y, T *out0, T *out1, int oh, int ow, int Pw, int oy, int ox, long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int x = (int)(i % ow), y = (int)(i / ow);
    float2 v = spec_ld<FMT>(Y, (long long)(y + oy) * Pw + (x + ox));
    out0[i] = (T)v.x;
    out1[i] = (T)v.y;
}

// ---------------------------------------------------------------- backward
// Filter gradient spectrum: G[f][c][p] = sum_k A_k[c] conj(DA_k[f]) + B_k[c] conj(DB_k[f]),
// where X_k = spectrum of the image pair (a + i b) and D_k of the delta pair (da + i db),
// separated per frequency with X(-p): A = (X + X*)/2, B = (X - X*)/2i (X* = conj X(-p)).
// Its inverse is sum_b in[b][c] (correlated with) delta[b][f].
__global__ void pointwise_gw_kernel(const float2 *X, const float2 *D, float2 *G, int pairs, int C, int F, int Ph,
                                    int Pw, long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long long P = (long long)Ph * Pw;
    const long long p = i % P;
    const long long r = i / P;
    const int c = (int)(r % C), f = (int)(r / C);
    const int v = (int)(p % Pw), u = (int)(p / Pw);
    const long long q = (long long)(u ? Ph - u : 0) * Pw + (v ? Pw - v : 0);
    float gr = 0.f, gi = 0.f;
    for (int k = 0; k < pairs; ++k) {
        const float2 xp = X[((long long)k * C + c) * P + p], xq = X[((long long)k * C + c) * P + q];
        const float2 dp = D[((long long)k * F + f) * P + p], dq = D[((long long)k * F + f) * P + q];
        // A = (xp + conj xq)/2, B = (xp - conj xq)/(2i)
        const float2 A = make_float2(0.5f * (xp.x + xq.x), 0.5f * (xp.y - xq.y));
        const float2 B = make_float2(0.5f * (xp.y + xq.y), -0.5f * (xp.x - xq.x));
        const float2 DA = make_float2(0.5f * (dp.x + dq.x), 0.5f * (dp.y - dq.y));
        const float2 DB = make_float2(0.5f * (dp.y + dq.y), -0.5f * (dp.x - dq.x));
        // A conj(DA) + B conj(DB)
        gr += A.x * DA.x + A.y * DA.y + B.x * DB.x + B.y * DB.y;
        gi += A.y * DA.x - A.x * DA.y + B.y * DB.x - B.x * DB.y;
    }
    G[i] = make_float2(gr, gi);
}

// Input gradient spectrum: R[k][c][p] = sum_f D_k[f][p] conj(H[f][c][p]) with H the spectrum
// of the 180-degree rotated filter; conj(H) is the spectrum of a real signal, so the image
// pair stays separable in the inverse.
__global__ void pointwise_gi_kernel(const float2 *D, const float2 *H, float2 *R, int C, int F, long long P,
                                    long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long long p = i % P;
    const long long r = i / P;
    const int c = (int)(r % C);
    const long long k = r / C;
    float re = 0.f, im = 0.f;
    for (int f = 0; f < F; ++f) {
        const float2 d = D[(k * F + f) * P + p];
        const float2 h = H[((long long)f * C + c) * P + p];
        re += d.x * h.x + d.y * h.y;
        im += d.y * h.x - d.x * h.y;
    }
    R[i] = make_float2(re, im);
}

// out[b][c][y][x] = scale * (Re or Im) Y[b/2][c][(y + oy) mod Ph][(x + ox) mod Pw]
template <typename T>
__global__ void crop_cyclic_kernel(const float2 *Y, T *out, int C, int oh, int ow, int Ph, int Pw, int oy, int ox,
                                   float scale, long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int x = (int)(i % ow);
    long long r = i / ow;
    int y = (int)(r % oh);
    r /= oh;
    int c = (int)(r % C);
    long long b = r / C;
    const int yy = (y + oy) % Ph, xx = (x + ox) % Pw;
    float2 v = Y[(((b >> 1) * C + c) * Ph + yy) * (long long)Pw + xx];
    out[i] = (T)((b & 1 ? v.y : v.x) * scale);
}

// out[n][y][x] = scale * Re Y[n][y][x] for the top-left oh x ow window of each of N planes
template <typename T>
__global__ void crop_real_kernel(const float2 *Y, T *out, int oh, int ow, int Ph, int Pw, float scale, long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int x = (int)(i % ow);
    long long r = i / ow;
    int y = (int)(r % oh);
    long long n = r / oh;
    out[i] = (T)(Y[(n * Ph + y) * (long long)Pw + x].x * scale);
}

static inline unsigned nblocks(long long n) { return (unsigned)((n + 255) / 256); }

template <int FMT>
static void launch_pack_pairs(const void *src, int f64, void *dst, int B, int C, int h, int w, int Ph, int Pw,
                              long long total, hipStream_t s)
{
    if (f64)
        hipLaunchKernelGGL((pack_pairs_kernel<double, FMT>), dim3(nblocks(total)), dim3(256), 0, s,
                           (const double *)src, dst, B, C, h, w, Ph, Pw, total);
    else
        hipLaunchKernelGGL((pack_pairs_kernel<float, FMT>), dim3(nblocks(total)), dim3(256), 0, s, (const float *)src,
                           dst, B, C, h, w, Ph, Pw, total);
}

hipError_t conv_pack_pairs(const void *src, int f64, void *dst, int fmt, int B, int C, int h, int w, int Ph, int Pw,
                           hipStream_t s)
{
    long long pairs = (B + 1) / 2, total = pairs * C * (long long)Ph * Pw;
    if (fmt == FMT_C16)
        launch_pack_pairs<FMT_C16>(src, f64, dst, B, C, h, w, Ph, Pw, total, s);
    else
        launch_pack_pairs<FMT_C32>(src, f64, dst, B, C, h, w, Ph, Pw, total, s);
    return hipGetLastError();
}

hipError_t conv_pack_filters(const void *src, int f64, float2 *dst, int FC, int fh, int fw, int Ph, int Pw, int rotate,
                             hipStream_t s)
{
    long long total = (long long)FC * Ph * Pw;
    if (f64)
        hipLaunchKernelGGL(pack_filters_kernel<double>, dim3(nblocks(total)), dim3(256), 0, s, (const double *)src, dst,
                           fh, fw, Ph, Pw, rotate, total);
    else
        hipLaunchKernelGGL(pack_filters_kernel<float>, dim3(nblocks(total)), dim3(256), 0, s, (const float *)src, dst,
                           fh, fw, Ph, Pw, rotate, total);
    return hipGetLastError();
}

hipError_t conv_pointwise(const void *X, const float2 *H, void *Y, int fmt, long long pairs, int C, int F, long long P,
                          hipStream_t s, int yfmt)
{
    long long total = pairs * F * P;
    if (yfmt < 0) yfmt = fmt;
    if (fmt == FMT_C16 && yfmt == FMT_C16)
        hipLaunchKernelGGL((pointwise_mac_kernel<FMT_C16, FMT_C16>), dim3(nblocks(total)), dim3(256), 0, s, X, H, Y, C,
                           F, P, total);
    else if (fmt == FMT_C16)
        hipLaunchKernelGGL((pointwise_mac_kernel<FMT_C16, FMT_C32>), dim3(nblocks(total)), dim3(256), 0, s, X, H, Y, C,
                           F, P, total);
    else
        hipLaunchKernelGGL((pointwise_mac_kernel<FMT_C32, FMT_C32>), dim3(nblocks(total)), dim3(256), 0, s, X, H, Y, C,
                           F, P, total);
    return hipGetLastError();
}

hipError_t conv_quantise(const float2 *R, void *Q, float *partial, float *scale, long long maps, long long P,
                         int nchunk, hipStream_t s)
{
    const long long blocks = maps * nchunk;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(quant_max_kernel, dim3((unsigned)blocks), dim3(256), 0, s, R, partial, P, nchunk);
    hipLaunchKernelGGL(quant_apply_kernel, dim3((unsigned)blocks), dim3(256), 0, s, R, (uint32_t *)Q, partial, scale, P,
                       nchunk);
    return hipGetLastError();
}

template <int FMT>
static hipError_t launch_crop(const void *Y, void *out, int f64, int B, int F, int oh, int ow, int Ph, int Pw, int oy,
                              int ox, float scale, hipStream_t s, const float *map_scale)
{
    const bool vec = !f64 && ow % 4 == 0 && ((uintptr_t)out & 15) == 0;
    const int xb = (int)nblocks(vec ? (ow + 3) / 4 : ow);
    const long long blocks = (long long)((B + 1) / 2) * F * oh * xb;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (f64)
        hipLaunchKernelGGL((crop_unpack_kernel<double, FMT, 1>), dim3((unsigned)blocks), dim3(256), 0, s, Y,
                           (double *)out, B, F, oh, ow, Ph, Pw, oy, ox, scale, xb, map_scale);
    else if (vec)
        hipLaunchKernelGGL((crop_unpack_kernel<float, FMT, 4>), dim3((unsigned)blocks), dim3(256), 0, s, Y,
                           (float *)out, B, F, oh, ow, Ph, Pw, oy, ox, scale, xb, map_scale);
    else
        hipLaunchKernelGGL((crop_unpack_kernel<float, FMT, 1>), dim3((unsigned)blocks), dim3(256), 0, s, Y,
                           (float *)out, B, F, oh, ow, Ph, Pw, oy, ox, scale, xb, map_scale);
    return hipGetLastError();
}

hipError_t conv_crop(const void *Y, int fmt, void *out, int f64, int B, int F, int oh, int ow, int Ph, int Pw, int oy,
                     int ox, float scale, hipStream_t s, const float *map_scale)
{
    if (fmt == FMT_C16) return launch_crop<FMT_C16>(Y, out, f64, B, F, oh, ow, Ph, Pw, oy, ox, scale, s, map_scale);
    return launch_crop<FMT_C32>(Y, out, f64, B, F, oh, ow, Ph, Pw, oy, ox, scale, s, map_scale);
}

hipError_t conv_pointwise_gw(const float2 *X, const float2 *D, float2 *G, int pairs, int C, int F, int Ph, int Pw,
                             hipStream_t s)
{
    long long total = (long long)F * C * Ph * Pw;
    hipLaunchKernelGGL(pointwise_gw_kernel, dim3(nblocks(total)), dim3(256), 0, s, X, D, G, pairs, C, F, Ph, Pw, total);
    return hipGetLastError();
}

hipError_t conv_pointwise_gi(const float2 *D, const float2 *H, float2 *R, long long pairs, int C, int F, long long P,
                             hipStream_t s)
{
    long long total = pairs * C * P;
    hipLaunchKernelGGL(pointwise_gi_kernel, dim3(nblocks(total)), dim3(256), 0, s, D, H, R, C, F, P, total);
    return hipGetLastError();
}

hipError_t conv_crop_cyclic(const float2 *Y, void *out, int f64, int B, int C, int oh, int ow, int Ph, int Pw, int oy,
                            int ox, float scale, hipStream_t s)
{
    long long total = (long long)B * C * oh * ow;
    if (f64)
        hipLaunchKernelGGL(crop_cyclic_kernel<double>, dim3(nblocks(total)), dim3(256), 0, s, Y, (double *)out, C, oh,
                           ow, Ph, Pw, oy, ox, scale, total);
    else
        hipLaunchKernelGGL(crop_cyclic_kernel<float>, dim3(nblocks(total)), dim3(256), 0, s, Y, (float *)out, C, oh,
                           ow, Ph, Pw, oy, ox, scale, total);
    return hipGetLastError();
}

hipError_t conv_crop_real(const float2 *Y, void *out, int f64, long long planes, int oh, int ow, int Ph, int Pw,
                          float scale, hipStream_t s)
{
    long long total = planes * oh * ow;
    if (f64)
        hipLaunchKernelGGL(crop_real_kernel<double>, dim3(nblocks(total)), dim3(256), 0, s, Y, (double *)out, oh, ow,
                           Ph, Pw, scale, total);
    else
        hipLaunchKernelGGL(crop_real_kernel<float>, dim3(nblocks(total)), dim3(256), 0, s, Y, (float *)out, oh, ow,
                           Ph, Pw, scale, total);
    return hipGetLastError();
}

template <int FMT>
static void launch_pack_two(const void *a, const void *b, int f64, void *dst, int h, int w, int Ph, int Pw,
                            long long total, hipStream_t s)
{
    if (f64)
        hipLaunchKernelGGL((pack_two_kernel<double, FMT>), dim3(nblocks(total)), dim3(256), 0, s, (const double *)a,
                           (const double *)b, dst, h, w, Ph, Pw, total);
    else
        hipLaunchKernelGGL((pack_two_kernel<float, FMT>), dim3(nblocks(total)), dim3(256), 0, s, (const float *)a,
                           (const float *)b, dst, h, w, Ph, Pw, total);
}

hipError_t conv_pack_two(const void *a, const void *b, int f64, void *dst, int fmt, int h, int w, int Ph, int Pw,
                         hipStream_t s)
{
    long long total = (long long)Ph * Pw;
    if (fmt == FMT_C16)
        launch_pack_two<FMT_C16>(a, b, f64, dst, h, w, Ph, Pw, total, s);
    else
        launch_pack_two<FMT_C32>(a, b, f64, dst, h, w, Ph, Pw, total, s);
    return hipGetLastError();
}

hipError_t conv_pointwise_two(const void *X, const float2 *H0, const float2 *H1, void *Y, int fmt, int Ph, int Pw,
                              hipStream_t s)
{
    if (fmt == FMT_C16)
        hipLaunchKernelGGL(pointwise_two_kernel<FMT_C16>, dim3(nblocks((long long)Ph * Pw)), dim3(256), 0, s, X, H0, H1,
                           Y, Ph, Pw);
    else
        hipLaunchKernelGGL(pointwise_two_kernel<FMT_C32>, dim3(nblocks((long long)Ph * Pw)), dim3(256), 0, s, X, H0, H1,
                           Y, Ph, Pw);
    return hipGetLastError();
}

template <int FMT>
static void launch_crop_two(const void *Y, void *out0, void *out1, int f64, int oh, int ow, int Pw, int oy, int ox,
                            long long total, hipStream_t s)
{
    if (f64)
        hipLaunchKernelGGL((crop_two_kernel<double, FMT>), dim3(nblocks(total)), dim3(256), 0, s, Y, (double *)out0,
                           (double *)out1, oh, ow, Pw, oy, ox, total);
    else
        hipLaunchKernelGGL((crop_two_kernel<float, FMT>), dim3(nblocks(total)), dim3(256), 0, s, Y, (float *)out0,
                           (float *)out1, oh, ow, Pw, oy, ox, total);
}

hipError_t conv_crop_two(const void *Y, int fmt, void *out0, void *out1, int f64, int oh, int ow, int Pw, int oy,
                         int ox, hipStream_t s)
{
    long long total = (long long)oh * ow;
    if (fmt == FMT_C16)
        launch_crop_two<FMT_C16>(Y, out0, out1, f64, oh, ow, Pw, oy, ox, total, s);
    else
        launch_crop_two<FMT_C32>(Y, out0, out1, f64, oh, ow, Pw, oy, ox, total, s);
    return hipGetLastError();
}

}  // namespace mfft
