/*
 * Self-test of the FP64 oracle, built with AddressSanitizer and UndefinedBehaviorSanitizer
 * (`make -C oracle asan`). TEST INFRASTRUCTURE ONLY: it calls every oracle entry point on
 * small, odd and edge sizes (n = 1, primes, mixed radices, single rows and columns, filters
 * as large as the image) so that any out-of-bounds access, use after free or undefined
 * arithmetic in the oracle is reported; the checks are identities the mathematics fixes
 * (FFT = brute-force DFT, inverse round trip, separability, gen_rev is a permutation,
 * 'valid' correlation of a delta, rounding of representable values and ties).
 * Exit status 0 when every check passes.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int failures = 0;
#define CHECK(c, ...)                                       \
    do {                                                    \
        if (!(c)) {                                         \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                   \
            fputc('\n', stderr);                            \
            ++failures;                                     \
        }                                                   \
    } while (0)

static double max_abs_diff(const double *a, const double *b, int64_t n)
{
    double m = 0;
    for (int64_t i = 0; i < n; ++i) m = fmax(m, fabs(a[i] - b[i]));
    return m;
}

static void test_fft_sizes(void)
{
    const int sizes[] = {1, 2, 3, 4, 5, 7, 8, 9, 11, 12, 13, 16, 17, 25, 30, 31, 49, 64, 97, 120, 128, 255, 256};
    for (size_t k = 0; k < sizeof sizes / sizeof sizes[0]; ++k) {
        const int n = sizes[k];
        const int batch = 3;
        double *x = malloc(sizeof(double) * 2 * n * batch), *y = malloc(sizeof(double) * 2 * n * batch),
               *z = malloc(sizeof(double) * 2 * n * batch), *w = malloc(sizeof(double) * 2 * n * batch);
        oracle_fill_uniform(x, 2 * n * batch, 100 + n, -1.0, 1.0);
        CHECK(oracle_fft(x, y, n, batch, -1, 1) == 0, "fft n=%d", n);
        oracle_dft(x, z, n, batch, -1);
        CHECK(max_abs_diff(y, z, 2 * n * batch) <= 1e-11 * n, "fft vs dft n=%d", n);
        CHECK(oracle_fft(y, w, n, batch, 1, 2) == 0, "inverse n=%d", n);
        for (int64_t i = 0; i < 2 * n * batch; ++i) w[i] /= n;
        CHECK(max_abs_diff(w, x, 2 * n * batch) <= 1e-12 * n, "round trip n=%d", n);
        /* the batch-loop harness shape gives the same transform */
        CHECK(oracle_fft_batch_loop(x, w, n, batch) == 0, "batch loop n=%d", n);
        CHECK(max_abs_diff(w, y, 2 * n * batch) <= 1e-12 * n, "batch loop vs fft n=%d", n);
        free(x);
        free(y);
        free(z);
        free(w);
    }
}

static void test_fft2d(void)
{
    const int shapes[][2] = {{1, 1}, {1, 7}, {5, 1}, {3, 4}, {8, 6}, {16, 17}};
    for (size_t k = 0; k < sizeof shapes / sizeof shapes[0]; ++k) {
        const int nx = shapes[k][0], ny = shapes[k][1], batch = 2, n = nx * ny;
        double *x = malloc(sizeof(double) * 2 * n * batch), *y = malloc(sizeof(double) * 2 * n * batch),
               *r = malloc(sizeof(double) * 2 * n * batch), *t = malloc(sizeof(double) * 2 * n * batch);
        oracle_fill_uniform(x, 2 * n * batch, 7 + k, -1.0, 1.0);
        CHECK(oracle_fft2d(x, y, nx, ny, batch, -1, 1) == 0, "fft2d %dx%d", nx, ny);
        /* separability: rows (dft along y), then columns (dft along x) through a transpose */
        oracle_dft(x, r, ny, (int64_t)nx * batch, -1);
        for (int b = 0; b < batch; ++b)
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j) {
                    const int64_t s = ((int64_t)b * nx + i) * ny + j, d = ((int64_t)b * ny + j) * nx + i;
                    t[2 * d] = r[2 * s];
                    t[2 * d + 1] = r[2 * s + 1];
                }
        oracle_dft(t, r, nx, (int64_t)ny * batch, -1);
        for (int b = 0; b < batch; ++b)
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j) {
                    const int64_t s = ((int64_t)b * ny + j) * nx + i, d = ((int64_t)b * nx + i) * ny + j;
                    t[2 * d] = r[2 * s];
                    t[2 * d + 1] = r[2 * s + 1];
                }
        CHECK(max_abs_diff(y, t, 2 * n * batch) <= 1e-11 * n, "fft2d vs separable dft %dx%d", nx, ny);
        free(x);
        free(y);
        free(r);
        free(t);
    }
}

static void test_gen_rev(void)
{
    const int lists[][6] = {{16, 16, 0}, {16, 16, 2, 0}, {16, 16, 4, 0}, {2, 0}, {16, 16, 2, 16, 16, 0},
                            {4, 2, 8, 0}, {16, 16, 16, 16, 0}};
    for (size_t k = 0; k < sizeof lists / sizeof lists[0]; ++k) {
        int m = 0, n = 1;
        while (lists[k][m]) n *= lists[k][m++];
        int *rev = malloc(sizeof(int) * n), *seen = calloc(n, sizeof(int));
        CHECK(oracle_gen_rev(n, lists[k], m, rev) == 0, "gen_rev list %zu", k);
        for (int i = 0; i < n; ++i) {
            CHECK(rev[i] >= 0 && rev[i] < n, "gen_rev range");
            if (rev[i] >= 0 && rev[i] < n) seen[rev[i]]++;
        }
        for (int i = 0; i < n; ++i) CHECK(seen[i] == 1, "gen_rev is a permutation (list %zu, %d)", k, i);
        CHECK(rev[0] == 0, "gen_rev fixes 0");
        free(rev);
        free(seen);
    }
    int bad[2] = {16, 16}, r1[8];
    CHECK(oracle_gen_rev(8, bad, 2, r1) == -1, "gen_rev rejects a list that does not multiply to n");
}

static void test_conv(void)
{
    /* 'valid' correlation with a unit impulse filter at (a, b) reads the signal shifted */
    const int sy = 9, sx = 7, fy = 3, fx = 4;
    double s[9 * 7], f[3 * 4] = {0}, out[(9 - 3 + 1) * (7 - 4 + 1)];
    oracle_fill_uniform(s, sy * sx, 3, 0.0, 1.0);
    f[1 * fx + 2] = 1.0;
    oracle_corr2d_valid(s, f, out, sy, sx, fy, fx);
    for (int i = 0; i < sy - fy + 1; ++i)
        for (int j = 0; j < sx - fx + 1; ++j)
            CHECK(out[i * (sx - fx + 1) + j] == s[(i + 1) * sx + j + 2], "corr impulse");
    /* filter as large as the image: one output, the dot product */
    double one;
    oracle_corr2d_valid(s, s, &one, sy, sx, sy, sx);
    double dot = 0;
    for (int i = 0; i < sy * sx; ++i) dot += s[i] * s[i];
    CHECK(fabs(one - dot) <= 1e-12 * dot, "corr full-size filter");
    /* 'full' convolution with a 1x1 filter scales; sizes 1 x n and n x 1 */
    double a[5], b1 = 2.0, full[5];
    oracle_fill_uniform(a, 5, 4, -1.0, 1.0);
    oracle_conv2d_full(a, &b1, full, 1, 5, 1, 1);
    for (int i = 0; i < 5; ++i) CHECK(full[i] == 2.0 * a[i], "conv full 1x1");
    double col[5 + 3 - 1], ker[3] = {1, 1, 1};
    oracle_conv2d_full(a, ker, col, 5, 1, 3, 1);
    CHECK(fabs(col[2] - (a[0] + a[1] + a[2])) < 1e-15, "conv full column");
    /* 'same' convolution: a delta image returns the flipped-centred filter; row windows */
    const int h = 6, w = 5, kh = 3, kw = 3;
    double img[6 * 5] = {0}, k9[9], same[6 * 5], part[6 * 5];
    img[2 * w + 2] = 1.0;
    oracle_fill_uniform(k9, 9, 5, -1.0, 1.0);
    oracle_conv2d_same(img, k9, same, h, w, kh, kw, 0, h, 1);
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
            CHECK(same[(2 + dy) * w + 2 + dx] == k9[(dy + 1) * kw + dx + 1], "conv same delta");
    memset(part, 0, sizeof part);
    oracle_conv2d_same(img, k9, part, h, w, kh, kw, 1, 3, 1);
    CHECK(max_abs_diff(part, same + 1 * w, 2 * w) == 0, "conv same row window (rows stored from out[0])");
}

static void test_rounding_and_errors(void)
{
    double v[] = {0.0, -0.0, 1.0, 65504.0, 65520.0, 1.0 + 1.0 / 2048, 1.0 + 3.0 / 2048, 6.103515625e-05,
                  5.960464477539063e-08, 2.98e-08, -1.5, 1e-300};
    const double want16[] = {0.0, -0.0, 1.0, 65504.0, INFINITY, 1.0, 1.0 + 4.0 / 2048, 6.103515625e-05,
                             5.960464477539063e-08, 0.0, -1.5, 0.0};
    double c[12];
    memcpy(c, v, sizeof v);
    oracle_round_fp16(c, 12);
    for (int i = 0; i < 12; ++i) CHECK(c[i] == want16[i], "round_fp16 %d: %g -> %g", i, v[i], c[i]);
    double bf[] = {1.0, 1.0 + 1.0 / 256, 1.0 + 3.0 / 256, 3.0e38};
    oracle_round_bf16(bf, 4);
    CHECK(bf[0] == 1.0 && bf[1] == 1.0 && bf[2] == 1.0 + 4.0 / 256, "round_bf16 ties to even");
    double t[4] = {1, 2, 3, 4}, s2[4] = {1, 2, 3, 5}, l2, mx;
    oracle_err_l2max(t, s2, 4, &l2, &mx);
    CHECK(fabs(l2 - 1.0 / sqrt(39.0)) < 1e-15 && fabs(mx - 0.2) < 1e-15, "err_l2max");
    CHECK(fabs(oracle_err_capped(t, s2, 2, 1) - 0.2 / 2) < 1e-15, "err_capped (sum over re and im / n / batch)");
    uint64_t r[5];
    oracle_splitmix64(1234567, r, 5);
    CHECK(r[0] == 6457827717110365317ULL && r[4] == 16408922859458223821ULL, "splitmix64 vector");
}

int main(void)
{
    test_fft_sizes();
    test_fft2d();
    test_gen_rev();
    test_conv();
    test_rounding_and_errors();
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("oracle self-test: all checks passed\n");
    return 0;
}
