// A C++ caller of the reference's conv-header entry points (tensorcore_fft_conv2d.h) that
// its own harness does not reach: TensorCoreFFTCreate / TensorCoreFFTExec
// (tensorcore_fft_conv2d_kernel.cu:1495-1648), fft_2d_c2c / ifft_2d_c2c
// (tensorcore_fft_conv2d.cpp:79-214) and the 10-argument conv2d_forward (.cpp:291-363).
// Inputs are read from and outputs written to raw little-endian float64 files, so a test
// can check them against an FP64 oracle.
//
//   conv_api_main tcfft <nx> <ny> <batch> <fft: 1 | -1> <in> <out>
//       in: [batch][nx][ny] complex of fp16-exact values, digit-reversed along both axes (the order
//       the reference's rev kernels produce); fp16 on the device, in place, natural-order
//       output, unnormalised. fft = 1 uses the stream overload of TensorCoreFFTExec,
//       fft = -1 the plain one.
//   conv_api_main c2c <x> <y> <batch> <streams> <fft: 1 | -1> <in> <out>
//       fft_2d_c2c / ifft_2d_c2c on [batch][y][x] complex (x contiguous), twice in a row
//       on one object (the second call reuses its plan and buffers); writes the second.
//   conv_api_main conv10 <S> <F> <sy> <sx> <fy> <fx> <signal> <filter> <out>
//       conv2d_forward(input, weight, out, S, F, sy, sx, fy, fx, GPU) on a
//       default-constructed object: 'valid' correlation, out [F][sy-fy+1][sx-fx+1].
#include <hip/hip_fp16.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tensorcore_fft_conv2d.h"

static std::vector<double> read_f64(const char *path, size_t n)
{
    std::vector<double> v(n);
    FILE *f = fopen(path, "rb");
    if (!f || fread(v.data(), sizeof(double), n, f) != n) {
        fprintf(stderr, "cannot read %zu doubles from %s\n", n, path);
        exit(2);
    }
    fclose(f);
    return v;
}

static void write_f64(const char *path, const std::vector<double> &v)
{
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) {
        fprintf(stderr, "cannot write %s\n", path);
        exit(2);
    }
    fclose(f);
}

static int run_tcfft(int nx, int ny, int batch, int fft, const char *in, const char *out)
{
    const size_t n = (size_t)nx * ny * batch * 2;
    std::vector<double> h = read_f64(in, n);
    std::vector<__half> hh(n);
    for (size_t i = 0; i < n; ++i) hh[i] = __float2half((float)h[i]);  // inputs are fp16-exact: no double rounding
    void *d = nullptr;
    if (hipMalloc(&d, n * sizeof(__half)) != hipSuccess) return 3;
    if (hipMemcpy(d, hh.data(), n * sizeof(__half), hipMemcpyHostToDevice) != hipSuccess) return 3;
    fft_conv2d_half layer;
    fftPlan plan;
    layer.TensorCoreFFTCreate(&plan, nx, ny, batch, fft, __AMPERE__);
    if (fft == 1) {
        hipStream_t s;
        if (hipStreamCreate(&s) != hipSuccess) return 3;
        layer.TensorCoreFFTExec(plan, d, s);
        if (hipStreamSynchronize(s) != hipSuccess) return 3;
        (void)hipStreamDestroy(s);
    } else {
        layer.TensorCoreFFTExec(plan, d);
        if (hipDeviceSynchronize() != hipSuccess) return 3;
    }
    if (hipMemcpy(hh.data(), d, n * sizeof(__half), hipMemcpyDeviceToHost) != hipSuccess) return 3;
    for (size_t i = 0; i < n; ++i) h[i] = __half2float(hh[i]);
    (void)hipFree(d);
    write_f64(out, h);
    printf("tcfft nx=%d ny=%d batch=%d fft=%d done\n", nx, ny, batch, fft);
    return 0;
}

static int run_c2c(int x, int y, int batch, int streams, int fft, const char *in, const char *out)
{
    const size_t n = (size_t)x * y * batch * 2;
    std::vector<double> h = read_f64(in, n), o(n, 0.0);
    fft_conv2d_half layer;
    for (int rep = 0; rep < 2; ++rep) {
        std::fill(o.begin(), o.end(), 0.0);
        if (fft == 1)
            layer.fft_2d_c2c(h.data(), o.data(), x, y, batch, __AMPERE__, streams);
        else
            layer.ifft_2d_c2c(h.data(), o.data(), x, y, batch, __AMPERE__, streams);
    }
    write_f64(out, o);
    printf("c2c x=%d y=%d batch=%d streams=%d fft=%d done\n", x, y, batch, streams, fft);
    return 0;
}

static int run_conv10(int S, int F, int sy, int sx, int fy, int fx, const char *sig, const char *filt, const char *out)
{
    std::vector<double> s = read_f64(sig, (size_t)S * sy * sx), w = read_f64(filt, (size_t)F * S * fy * fx);
    std::vector<double> o((size_t)F * (sy - fy + 1) * (sx - fx + 1));
    fft_conv2d_half layer;
    layer.conv2d_forward(s.data(), w.data(), o.data(), S, F, sy, sx, fy, fx, __AMPERE__);
    write_f64(out, o);
    printf("conv10 S=%d F=%d %dx%d filter %dx%d pad %dx%d done\n", S, F, sy, sx, fy, fx, layer.padded_y,
           layer.padded_x);
    return 0;
}

int main(int argc, char **argv)
{
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "tcfft" && argc == 8)
            return run_tcfft(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), argv[6], argv[7]);
        if (mode == "c2c" && argc == 9)
            return run_c2c(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]), argv[7], argv[8]);
        if (mode == "conv10" && argc == 11)
            return run_conv10(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]), atoi(argv[7]),
                              argv[8], argv[9], argv[10]);
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    fprintf(stderr, "usage: conv_api_main tcfft|c2c|conv10 ...\n");
    return 2;
}
