// A C++ caller of the reference's conv operator interface (tensorcore_fft_conv2d.h),
// following main.cpp's accuracy_conv (main.cpp:286-317): one S-channel signal, F filters,
// 'valid' cross-correlation, checked against the direct FP64 correlation
// (naive_fft_conv2d_ss_sf_valid, main.cpp:97-116) with the reference's capped metric
// (get_error_conv, main.cpp:136-148) and rel-L2.
//   usage: conv_main [s] [f] [filters]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tensorcore_fft_conv2d.h"

int main(int argc, char **argv)
{
    int s = argc > 1 ? atoi(argv[1]) : 256, f = argc > 2 ? atoi(argv[2]) : 13, F = argc > 3 ? atoi(argv[3]) : 16;
    const int C = 1;
    std::vector<double> signal((size_t)s * s * C), filter((size_t)f * f * F * C);
    srand(42);
    for (auto &v : signal) v = 0.1 * (rand() % 10);
    for (auto &v : filter) v = 0.1 * (rand() % 10);
    int o = s - f + 1;
    std::vector<double> out((size_t)o * o * F), ref(out.size());
    fft_conv2d_half layer(signal.data(), filter.data(), C, F, s, s, f, f, __AMPERE__);
    layer.conv2d_forward(out.data());
    double worst_l2 = 0, capped = 0;
    for (int k = 0; k < F; ++k) {
        double num = 0, den = 0;
        for (int i = 0; i < o; ++i)
            for (int j = 0; j < o; ++j) {
                double acc = 0;
                for (int a = 0; a < f; ++a)
                    for (int b = 0; b < f; ++b) acc += filter[(size_t)k * f * f + a * f + b] * signal[(size_t)(i + a) * s + j + b];
                double t = out[(size_t)k * o * o + (size_t)i * o + j];
                num += (t - acc) * (t - acc);
                den += acc * acc;
                double r = std::fabs((t - acc) / acc);
                capped += r < 1 ? r : 1;
            }
        worst_l2 = std::fmax(worst_l2, std::sqrt(num / den));
    }
    printf("conv s=%d f=%d F=%d pad=%d capped=%.3e rel_l2=%.3e\n", s, f, F, layer.padded_x, capped / ((double)o * o * F),
           worst_l2);
    // backward (tensorcore_fft_conv2d.cpp:364-442): dL/dout = the forward output itself;
    // filter gradient of filter 0 against the direct 'valid' correlation signal * dL/dout
    layer.conv2d_backward(out.data());
    double num = 0, den = 0;
    for (int a = 0; a < f; ++a)
        for (int b = 0; b < f; ++b) {
            double acc = 0;
            for (int i = 0; i < o; ++i)
                for (int j = 0; j < o; ++j) acc += signal[(size_t)(i + a) * s + j + b] * out[(size_t)i * o + j];
            double t = layer.delta_for_filter[(size_t)a * f + b];
            num += (t - acc) * (t - acc);
            den += acc * acc;
        }
    const double bw_l2 = std::sqrt(num / den);
    printf("conv_backward filter0 rel_l2=%.3e\n", bw_l2);
    return worst_l2 <= 1e-4 && bw_l2 <= 1e-4 ? 0 : 1;
}
