// A C++ program written against the reference's interface (the calls main.cpp makes:
// setup/doit/finalize, setup_2d/doit_2d/finalize_2d, tcfftCreate/tcfftExec), linked to
// the MI355X library instead. Prints the reference's capped error metric
// (main.cpp:118-134) and rel-L2 against an FP64 DFT computed here.
//   usage: compat_main <n> <batch> [nx ny]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tcfft_half.h"
#include "tcfft_half_2d.h"

static void dft64(const double *x, double *y, int n)
{
    for (int k = 0; k < n; ++k) {
        double sr = 0, si = 0;
        for (int j = 0; j < n; ++j) {
            double a = -2.0 * M_PI * (double)((long long)j * k % n) / n;
            sr += x[2 * j] * cos(a) - x[2 * j + 1] * sin(a);
            si += x[2 * j] * sin(a) + x[2 * j + 1] * cos(a);
        }
        y[2 * k] = sr;
        y[2 * k + 1] = si;
    }
}

static double capped(const double *t, const double *s, long long n)
{
    double e = 0;
    for (long long j = 0; j < 2 * n; ++j) {
        double r = fabs((t[j] - s[j]) / s[j]);
        e += r < 1.0 ? r : 1.0;
    }
    return e / n;
}

static double rel_l2(const double *t, const double *s, long long m)
{
    double a = 0, b = 0;
    for (long long j = 0; j < m; ++j) {
        a += (t[j] - s[j]) * (t[j] - s[j]);
        b += s[j] * s[j];
    }
    return sqrt(a / b);
}

int main(int argc, char **argv)
{
    int n = argc > 1 ? atoi(argv[1]) : 256, batch = argc > 2 ? atoi(argv[2]) : 4;
    int nx = argc > 4 ? atoi(argv[3]) : 16, ny = argc > 4 ? atoi(argv[4]) : 32;
    std::vector<double> data((size_t)2 * n * batch), res(data.size()), ref(data.size());
    srand(42);
    for (auto &v : data) v = 0.1 * (rand() % 10);  // the reference's generator (main.cpp:32-53)
    for (int b = 0; b < batch; ++b) dft64(&data[2 * (size_t)n * b], &ref[2 * (size_t)n * b], n);

    // harness path (natural order, main.cpp:171-174)
    setup(data.data(), n, batch);
    doit(1);
    finalize(res.data(), batch);
    printf("harness1d n=%d batch=%d err=%d capped=%.3e rel_l2=%.3e\n", n, batch, tcfftGetLastError(),
           capped(res.data(), ref.data(), (long long)n * batch), rel_l2(res.data(), ref.data(), res.size()));

    // plan/exec path: host digit reversal with the plan's radices, as setup() does
    tcfftHandle plan;
    tcfftCreate(&plan, n, batch);
    std::vector<int> rev(n), a(n), bb(n);
    for (int i = 0; i < n; ++i) a[i] = i;
    int now = n;
    for (int r = plan.n_radices - 1; r >= 0; --r) {  // gen_rev (tcfft_half.cu:1351-1374)
        int rr = plan.radices[r], sub = now / rr;
        for (int j = 0; j < n; j += now)
            for (int k = 0; k < rr; ++k)
                for (int l = 0; l < sub; ++l) bb[j + l + k * sub] = a[j + l * rr + k];
        now = sub;
        a.swap(bb);
    }
    rev = a;
    std::vector<_Float16> h((size_t)2 * n * batch);
    for (int b = 0; b < batch; ++b)
        for (int i = 0; i < n; ++i) {
            h[2 * ((size_t)n * b + i)] = (_Float16)data[2 * ((size_t)n * b + rev[i])];
            h[2 * ((size_t)n * b + i) + 1] = (_Float16)data[2 * ((size_t)n * b + rev[i]) + 1];
        }
    void *d = nullptr;
    if (hipMalloc(&d, h.size() * 2) != hipSuccess) return 2;
    if (hipMemcpy(d, h.data(), h.size() * 2, hipMemcpyHostToDevice) != hipSuccess) return 2;
    tcfftExec(plan, d);
    if (hipMemcpy(h.data(), d, h.size() * 2, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    for (size_t i = 0; i < h.size(); ++i) res[i] = (double)h[i];
    printf("exec1d n=%d batch=%d err=%d capped=%.3e rel_l2=%.3e passes=%d\n", n, batch, tcfftGetLastError(),
           capped(res.data(), ref.data(), (long long)n * batch), rel_l2(res.data(), ref.data(), res.size()),
           plan.n_mergings);
    tcfftDestroy(&plan);
    (void)hipFree(d);

    // 2D harness: FP64 2D DFT by rows then columns
    std::vector<double> d2((size_t)2 * nx * ny), r2(d2.size()), f2(d2.size()), tmp(2 * (size_t)(nx > ny ? nx : ny)),
        o(tmp.size());
    for (auto &v : d2) v = 0.1 * (rand() % 10);
    for (int x = 0; x < nx; ++x) dft64(&d2[2 * (size_t)x * ny], &f2[2 * (size_t)x * ny], ny);
    for (int y = 0; y < ny; ++y) {
        for (int x = 0; x < nx; ++x) {
            tmp[2 * x] = f2[2 * ((size_t)x * ny + y)];
            tmp[2 * x + 1] = f2[2 * ((size_t)x * ny + y) + 1];
        }
        dft64(tmp.data(), o.data(), nx);
        for (int x = 0; x < nx; ++x) {
            f2[2 * ((size_t)x * ny + y)] = o[2 * x];
            f2[2 * ((size_t)x * ny + y) + 1] = o[2 * x + 1];
        }
    }
    setup_2d(d2.data(), nx, ny, 1);
    doit_2d(1);
    finalize_2d(r2.data(), 1);
    printf("harness2d %dx%d err=%d capped=%.3e rel_l2=%.3e\n", nx, ny, tcfftGetLastError(),
           capped(r2.data(), f2.data(), (long long)nx * ny), rel_l2(r2.data(), f2.data(), r2.size()));
    return 0;
}
