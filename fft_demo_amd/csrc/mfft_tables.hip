// Twiddle and DFT-matrix tables, generated on the GPU in FP64 and rounded once to the
// storage precision (fp32 twiddles; fp16 hi+lo and bf16 DFT16 operands).
//
// The reference builds its 16x16 DFT matrix on the host with cosf/sinf rounded to half
// (tcfft_half.cu:1332-1340) and raises twiddles to powers by repeated fp16/fp32 complex
// multiplies inside the kernels (tcfft_half.cu:131-141, 306-321). Here every entry is a
// direct FP64 evaluation of exp(s*2*pi*i*e/n) with an exact integer phase e mod n.
#include "mfft_common.h"

namespace mfft {

// out[i] = scale * exp(sign * 2*pi*i * ((i * mult) mod n) / n), i in [0, count); the
// scale (a power of two) is applied in FP64 before the single rounding to fp32.
__global__ void gen_twiddle_kernel(float2 *out, long long n, long long count, long long mult, int sign, double scale)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    long long e = (long long)(((__int128)i * mult) % n);
    double sn, cs;
    sincospi(2.0 * (double)e / (double)n, &sn, &cs);
    out[i] = make_float2((float)(scale * cs), (float)(scale * sign * sn));
}

// DFT16 operands for lane l: rows m = l % 16 of the K-folded real matrix, K entries
// 8*(l/16)..+7 = complex inputs j = 4*(l/16) + v/2, component v % 2.
//   real output row: [cos, -s*sin] ; imaginary output row: [s*sin, cos]
// (theta = 2*pi*m*j/16), all multiplied by 16 (exact) so the fp16 lo parts stay normal.
__global__ void gen_dft_ops_kernel(DftOps *ops, int sign)
{
    int l = threadIdx.x;
    if (l >= 64) return;
    int g = l >> 4, m = l & 15;
    for (int v = 0; v < 8; ++v) {
        int j = 4 * g + (v >> 1);
        int comp = v & 1;
        double sn, cs;
        sincospi((double)((m * j) % 16) / 8.0, &sn, &cs);
        double rr = comp == 0 ? cs : -sign * sn;
        double ri = comp == 0 ? sign * sn : cs;
        rr *= 16.0;
        ri *= 16.0;
        _Float16 hr = (_Float16)rr, hi = (_Float16)ri;
        ops->hr[l][v] = hr;
        ops->hi_[l][v] = hi;
        ops->lr[l][v] = (_Float16)(rr - (double)hr);
        ops->li[l][v] = (_Float16)(ri - (double)hi);
        ops->br[l][v] = (__bf16)(float)rr;
        ops->bi[l][v] = (__bf16)(float)ri;
    }
}

// cos(pi/8), sin(pi/8), cos(pi/4) in FP64, rounded once to fp32: the constants of the VALU
// radix-8 and radix-16 butterflies (TileArgs::w16).
__global__ void gen_w16_kernel(float *out)
{
    if (threadIdx.x != 0) return;
    double sn, cs;
    sincospi(1.0 / 8.0, &sn, &cs);
    out[0] = (float)cs;
    out[1] = (float)sn;
    out[2] = (float)cospi(1.0 / 4.0);
}

hipError_t gen_w16(float *host3)
{
    float *d = nullptr;
    hipError_t e = hipMalloc(&d, 3 * sizeof(float));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gen_w16_kernel, dim3(1), dim3(64), 0, 0, d);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(host3, d, 3 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e;
}

hipError_t gen_twiddle(float2 *out, long long n, long long count, long long mult, int sign, hipStream_t s,
                       double scale)
{
    if (count <= 0) return hipSuccess;
    unsigned blocks = (unsigned)((count + 255) / 256);
    hipLaunchKernelGGL(gen_twiddle_kernel, dim3(blocks), dim3(256), 0, s, out, n, count, mult, sign, scale);
    return hipGetLastError();
}

hipError_t gen_dft_ops(DftOps *ops, int sign, hipStream_t s)
{
    hipLaunchKernelGGL(gen_dft_ops_kernel, dim3(1), dim3(64), 0, s, ops, sign);
    return hipGetLastError();
}

}  // namespace mfft
