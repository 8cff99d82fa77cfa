// On-device staging: format conversion (fp64 / fp32 / fp16 / bf16 complex) with an
// optional mixed-radix digit reversal and scale, in one pass over HBM.
//
// The reference stages on the host: setup() converts FP64 to half and applies the
// digit-reversal permutation in a host loop before the H2D copy (tcfft_half.cu:1385-1393,
// tcfft_half_2d.cu:1264-1271), finalize() converts back after the D2H copy
// (tcfft_half.cu:1396-1406); the conv operator does the same on the device in its rev /
// half2double kernels (tensorcore_fft_conv2d_kernel.cu:1277-1285, 1407-1414).
#include "mfft_common.h"

namespace mfft {

__device__ __forceinline__ float2 ld_any(const void *p, long long i, int fmt)
{
    switch (fmt) {
    case FMT_C32: return reinterpret_cast<const float2 *>(p)[i];
    case 3: {
        double2 d = reinterpret_cast<const double2 *>(p)[i];
        return make_float2((float)d.x, (float)d.y);
    }
    default: {
        uint32_t w = reinterpret_cast<const uint32_t *>(p)[i];
        if (fmt == FMT_C16)
            return make_float2((float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffff)),
                               (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)));
        return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
    }
    }
}

// FP64 source values are rounded once, directly to the destination format.
__global__ void stage_kernel(const void *src, int sfmt, void *dst, int dfmt, long long n_inner, long long total,
                             const int *perm_in, int pn, const int *perm_y, int py, double scale)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    // destination index i; source index = permuted position (gather)
    long long b = i / n_inner, r = i - b * n_inner;
    long long sr = r;
    if (perm_y) {
        // 2D: r = x*py + y ; source (perm_in[x], perm_y[y])
        long long x = r / py, y = r - x * py;
        sr = (long long)(perm_in ? perm_in[x] : x) * py + perm_y[y];
    } else if (perm_in) {
        sr = perm_in[r];
    }
    long long si = b * n_inner + sr;
    if (dfmt == 3) {
        double2 v;
        if (sfmt == 3) {
            v = reinterpret_cast<const double2 *>(src)[si];
        } else {
            float2 f = ld_any(src, si, sfmt);
            v = make_double2(f.x, f.y);
        }
        v.x *= scale;
        v.y *= scale;
        reinterpret_cast<double2 *>(dst)[i] = v;
        return;
    }
    double re, im;
    if (sfmt == 3) {
        double2 d = reinterpret_cast<const double2 *>(src)[si];
        re = d.x;
        im = d.y;
    } else {
        float2 f = ld_any(src, si, sfmt);
        re = f.x;
        im = f.y;
    }
    re *= scale;
    im *= scale;
    if (dfmt == FMT_C32) {
        reinterpret_cast<float2 *>(dst)[i] = make_float2((float)re, (float)im);
    } else if (dfmt == FMT_C16) {
        uint32_t w = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)re) |
                     ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)im) << 16);
        reinterpret_cast<uint32_t *>(dst)[i] = w;
    } else {
        uint32_t w = (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)re) |
                     ((uint32_t)__builtin_bit_cast(uint16_t, (__bf16)im) << 16);
        reinterpret_cast<uint32_t *>(dst)[i] = w;
    }
}

hipError_t stage_convert(const void *src, int sfmt, void *dst, int dfmt, long long n_inner, long long batch,
                         const int *perm_x, int nx, const int *perm_y, int ny, double scale, hipStream_t s)
{
    long long total = n_inner * batch;
    if (total <= 0) return hipSuccess;
    unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(stage_kernel, dim3(blocks), dim3(256), 0, s, src, sfmt, dst, dfmt, n_inner, total, perm_x,
                       nx, perm_y, ny, scale);
    return hipGetLastError();
}

}  // namespace mfft
