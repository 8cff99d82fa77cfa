/*
 * FP64 CPU oracle for the matrix-core FFT / FFT-conv2d path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py as the checker. The product library never links or calls it.
 *
 * Reference behaviour restated here (yuelingfan/fft_demo):
 *   gen_rev ................ tcfft_half.cu:1351-1374 (= tcfft_half_2d.cu:1227-1250)
 *   FFT convention ......... main.cpp:68-94 (FFTW_FORWARD, unnormalised, interleaved)
 *   inverse sign ........... tensorcore_fft_conv2d_kernel.cu:47-82, 1634-1643
 *   'valid' correlation .... main.cpp:97-116
 *   capped error metric .... main.cpp:118-134
 *   batch-loop harness ..... main.cpp:68-80
 * The FFT itself is a textbook mixed-radix Stockham (decimation in frequency) in FP64;
 * FFTW, which the reference calls, is absent from this image, so the FFT is pinned by
 * brute-force DFTs, closed forms and an independent library in the tests.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO_PI 6.283185307179586476925286766559

/* ------------------------------------------------------------------ gen_rev */
/* Restates gen_rev (tcfft_half.cu:1351-1374): starting from the identity, for each
 * radix from last to first, every block of length now_N is re-ordered so that element
 * (l * r + k) moves to (l + k * now_N / r); then now_N /= r. */
int oracle_gen_rev(int n, const int *radices, int n_radices, int *rev)
{
    long prod = 1;
    for (int i = 0; i < n_radices; ++i) prod *= radices[i];
    if (prod != n || n <= 0) return -1;
    int *a = (int *)malloc(sizeof(int) * (size_t)n);
    int *b = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) a[i] = i;
    int now_n = n;
    for (int i = n_radices - 1; i >= 0; --i) {
        int r = radices[i];
        int sub = now_n / r;
        for (int j = 0; j < n; j += now_n)
            for (int k = 0; k < r; ++k)
                for (int l = 0; l < sub; ++l) b[j + l + k * sub] = a[j + l * r + k];
        now_n = sub;
        int *t = a;
        a = b;
        b = t;
    }
    memcpy(rev, a, sizeof(int) * (size_t)n);
    free(a);
    free(b);
    return 0;
}

/* ------------------------------------------------------------------ data */
static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void oracle_splitmix64(uint64_t seed, uint64_t *out, int64_t count)
{
    uint64_t s = seed;
    for (int64_t i = 0; i < count; ++i) out[i] = splitmix64(&s);
}

void oracle_fill_uniform(double *dst, int64_t count, uint64_t seed, double lo, double hi)
{
    uint64_t s = seed;
    for (int64_t i = 0; i < count; ++i) {
        double u = (double)(splitmix64(&s) >> 11) * (1.0 / 9007199254740992.0);
        dst[i] = lo + (hi - lo) * u;
    }
}

/* Round to a binary format with `mant` explicit fraction bits, minimum normal exponent
 * emin and maximum finite value maxv; ties to even (nearbyint under the default mode). */
static double round_to_format(double x, int mant, int emin, double maxv)
{
    if (x == 0.0 || !isfinite(x)) return x;
    int e;
    frexp(x, &e); /* |x| = m * 2^e, 0.5 <= m < 1, so the unit exponent is e - 1 */
    int ue = e - 1;
    if (ue < emin) ue = emin;
    double q = ldexp(1.0, ue - mant);
    double r = nearbyint(x / q) * q;
    if (fabs(r) > maxv) return x > 0 ? INFINITY : -INFINITY;
    return r;
}

void oracle_round_fp16(double *data, int64_t count)
{
    for (int64_t i = 0; i < count; ++i) data[i] = round_to_format(data[i], 10, -14, 65504.0);
}

void oracle_round_bf16(double *data, int64_t count)
{
    const double maxv = ldexp(255.0 / 128.0, 127);
    for (int64_t i = 0; i < count; ++i) data[i] = round_to_format(data[i], 7, -126, maxv);
}

/* ------------------------------------------------------------------ DFT */
void oracle_dft(const double *in, double *out, int64_t n, int64_t batch, int sign)
{
    double *wr = (double *)malloc(sizeof(double) * (size_t)n);
    double *wi = (double *)malloc(sizeof(double) * (size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        wr[k] = cos(TWO_PI * (double)k / (double)n);
        wi[k] = sign * sin(TWO_PI * (double)k / (double)n);
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t b = 0; b < batch; ++b)
        for (int64_t k = 0; k < n; ++k) {
            const double *x = in + 2 * n * b;
            double sr = 0.0, si = 0.0;
            for (int64_t j = 0; j < n; ++j) {
                int64_t e = (j * k) % n;
                double xr = x[2 * j], xi = x[2 * j + 1];
                sr += xr * wr[e] - xi * wi[e];
                si += xr * wi[e] + xi * wr[e];
            }
            out[2 * (n * b + k)] = sr;
            out[2 * (n * b + k) + 1] = si;
        }
    free(wr);
    free(wi);
}

/* ------------------------------------------------------------------ FFT */
typedef struct {
    int64_t n;
    int sign;
    int nf;
    int fac[64];
    double *w; /* w[2e], w[2e+1] = cos, sign*sin of 2*pi*e/n, e in [0, n) */
} fft_plan64;

static int factorize(int64_t n, int *fac)
{
    int nf = 0;
    while (n % 4 == 0) { fac[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { fac[nf++] = 2; n /= 2; }
    for (int64_t p = 3; p * p <= n; p += 2)
        while (n % p == 0) { fac[nf++] = (int)p; n /= p; }
    if (n > 1) fac[nf++] = (int)n;
    return nf;
}

static int plan64_init(fft_plan64 *p, int64_t n, int sign)
{
    p->n = n;
    p->sign = sign;
    p->nf = factorize(n, p->fac);
    p->w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    if (!p->w) return -1;
    for (int64_t e = 0; e < n; ++e) {
        /* reduce the phase to the first octant-free form e/n in [0, 1) exactly */
        double ang = TWO_PI * (double)e / (double)n;
        p->w[2 * e] = cos(ang);
        p->w[2 * e + 1] = sign * sin(ang);
    }
    return 0;
}

/* One Stockham DIF stage: length-nc sub-transforms at stride s, radix r.
 * y[q + s*(r*p + t)] = W_nc^(p*t) * sum_j x[q + s*(p + j*m)] * W_r^(j*t), m = nc / r. */
static void stockham_stage(const fft_plan64 *pl, const double *x, double *y, int64_t nc, int64_t s,
                           int r, double *tmp)
{
    const int64_t n = pl->n;
    const int64_t m = nc / r;
    const int64_t wstep_r = n / r;   /* W_r^1 = W_n^(n/r) */
    const int64_t wstep_nc = n / nc; /* W_nc^1 = W_n^(n/nc) */
    for (int64_t p = 0; p < m; ++p) {
        for (int64_t q = 0; q < s; ++q) {
            double *a = tmp;          /* r inputs */
            double *o = tmp + 2 * r;  /* r outputs */
            for (int j = 0; j < r; ++j) {
                a[2 * j] = x[2 * (q + s * (p + j * m))];
                a[2 * j + 1] = x[2 * (q + s * (p + j * m)) + 1];
            }
            if (r == 2) {
                o[0] = a[0] + a[2];
                o[1] = a[1] + a[3];
                o[2] = a[0] - a[2];
                o[3] = a[1] - a[3];
            } else if (r == 4) {
                /* W_4 = sign * i */
                double s0r = a[0] + a[4], s0i = a[1] + a[5];
                double d0r = a[0] - a[4], d0i = a[1] - a[5];
                double s1r = a[2] + a[6], s1i = a[3] + a[7];
                double d1r = a[2] - a[6], d1i = a[3] - a[7];
                /* rotate d1 by W_4^1 = (0, sign) : (d1r, d1i) * (0 + i*sign) */
                double rr = -pl->sign * d1i, ri = pl->sign * d1r;
                o[0] = s0r + s1r; o[1] = s0i + s1i;
                o[2] = d0r + rr;  o[3] = d0i + ri;
                o[4] = s0r - s1r; o[5] = s0i - s1i;
                o[6] = d0r - rr;  o[7] = d0i - ri;
            } else {
                for (int t = 0; t < r; ++t) {
                    double sr = 0.0, si = 0.0;
                    for (int j = 0; j < r; ++j) {
                        int64_t e = ((int64_t)(j * t) % r) * wstep_r;
                        double wr = pl->w[2 * e], wi = pl->w[2 * e + 1];
                        sr += a[2 * j] * wr - a[2 * j + 1] * wi;
                        si += a[2 * j] * wi + a[2 * j + 1] * wr;
                    }
                    o[2 * t] = sr;
                    o[2 * t + 1] = si;
                }
            }
            for (int t = 0; t < r; ++t) {
                int64_t e = ((p * t) % nc) * wstep_nc;
                double wr = pl->w[2 * e], wi = pl->w[2 * e + 1];
                double vr = o[2 * t], vi = o[2 * t + 1];
                int64_t dst = q + s * (r * p + t);
                y[2 * dst] = vr * wr - vi * wi;
                y[2 * dst + 1] = vr * wi + vi * wr;
            }
        }
    }
}

/* Transform one sequence: x (input, clobbered) -> out. work has 2n doubles. */
static void fft_one(const fft_plan64 *pl, double *x, double *work, double *out, double *tmp)
{
    double *a = x, *b = work;
    int64_t nc = pl->n, s = 1;
    for (int f = 0; f < pl->nf; ++f) {
        int r = pl->fac[f];
        stockham_stage(pl, a, b, nc, s, r, tmp);
        nc /= r;
        s *= r;
        double *t = a;
        a = b;
        b = t;
    }
    if (a != out) memcpy(out, a, sizeof(double) * 2 * (size_t)pl->n);
}

static int max_radix(const fft_plan64 *pl)
{
    int m = 4;
    for (int f = 0; f < pl->nf; ++f)
        if (pl->fac[f] > m) m = pl->fac[f];
    return m;
}

int oracle_fft(const double *in, double *out, int64_t n, int64_t batch, int sign, int threads)
{
    if (n < 1 || batch < 0) return -1;
    fft_plan64 pl;
    if (plan64_init(&pl, n, sign)) return -1;
    int mr = max_radix(&pl);
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    int nt = 1;
    (void)threads;
#endif
#pragma omp parallel num_threads(nt)
    {
        double *x = (double *)malloc(sizeof(double) * 2 * (size_t)n);
        double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
        double *tmp = (double *)malloc(sizeof(double) * 4 * (size_t)mr);
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < batch; ++b) {
            memcpy(x, in + 2 * n * b, sizeof(double) * 2 * (size_t)n);
            fft_one(&pl, x, w, out + 2 * n * b, tmp);
        }
        free(x);
        free(w);
        free(tmp);
    }
    free(pl.w);
    return 0;
}

int oracle_fft_batch_loop(const double *data, double *result, int64_t n, int64_t batch)
{
    fft_plan64 pl;
    if (plan64_init(&pl, n, -1)) return -1;
    double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *tmp = (double *)malloc(sizeof(double) * 4 * (size_t)max_radix(&pl));
    for (int64_t b = 0; b < batch; ++b) {
        memcpy(buf, data + 2 * n * b, sizeof(double) * 2 * (size_t)n);
        fft_one(&pl, buf, w, buf, tmp);
        memcpy(result + 2 * n * b, buf, sizeof(double) * 2 * (size_t)n);
    }
    free(buf);
    free(w);
    free(tmp);
    free(pl.w);
    return 0;
}

int oracle_fft2d(const double *in, double *out, int64_t nx, int64_t ny, int64_t batch, int sign,
                 int threads)
{
    if (nx < 1 || ny < 1) return -1;
    /* rows: nx*batch contiguous transforms of length ny */
    if (oracle_fft(in, out, ny, nx * batch, sign, threads)) return -1;
    fft_plan64 pl;
    if (plan64_init(&pl, nx, sign)) return -1;
    int mr = max_radix(&pl);
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    int nt = 1;
#endif
#pragma omp parallel num_threads(nt)
    {
        double *x = (double *)malloc(sizeof(double) * 2 * (size_t)nx);
        double *w = (double *)malloc(sizeof(double) * 2 * (size_t)nx);
        double *o = (double *)malloc(sizeof(double) * 2 * (size_t)nx);
        double *tmp = (double *)malloc(sizeof(double) * 4 * (size_t)mr);
#pragma omp for collapse(2) schedule(static)
        for (int64_t b = 0; b < batch; ++b)
            for (int64_t yy = 0; yy < ny; ++yy) {
                double *base = out + 2 * (b * nx * ny + yy);
                for (int64_t xx = 0; xx < nx; ++xx) {
                    x[2 * xx] = base[2 * xx * ny];
                    x[2 * xx + 1] = base[2 * xx * ny + 1];
                }
                fft_one(&pl, x, w, o, tmp);
                for (int64_t xx = 0; xx < nx; ++xx) {
                    base[2 * xx * ny] = o[2 * xx];
                    base[2 * xx * ny + 1] = o[2 * xx + 1];
                }
            }
        free(x);
        free(w);
        free(o);
        free(tmp);
    }
    free(pl.w);
    return 0;
}

/* ------------------------------------------------------------------ conv */
void oracle_corr2d_valid(const double *signal, const double *filter, double *out, int sy, int sx,
                         int fy, int fx)
{
    int ox = sx - fx + 1, oy = sy - fy + 1;
    for (int i = 0; i < oy; i++)
        for (int j = 0; j < ox; j++) {
            double acc = 0.0;
            for (int k = 0; k < fy; k++)
                for (int l = 0; l < fx; l++) acc += filter[k * fx + l] * signal[(i + k) * sx + j + l];
            out[i * ox + j] = acc;
        }
}

void oracle_conv2d_full(const double *a, const double *b, double *out, int ah, int aw, int bh, int bw)
{
    const int oh = ah + bh - 1, ow = aw + bw - 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            double acc = 0.0;
            for (int i = 0; i < bh; ++i) {
                const int ay = y - i;
                if (ay < 0 || ay >= ah) continue;
                for (int j = 0; j < bw; ++j) {
                    const int ax = x - j;
                    if (ax >= 0 && ax < aw) acc += a[ay * aw + ax] * b[i * bw + j];
                }
            }
            out[y * ow + x] = acc;
        }
}

void oracle_conv2d_same(const double *image, const double *filter, double *out, int h, int w,
                        int fh, int fw, int row0, int row1, int threads)
{
    const int oh = (fh - 1) / 2, ow = (fw - 1) / 2;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    int nt = 1;
    (void)threads;
#endif
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4)
    for (int i = row0; i < row1; ++i)
        for (int j = 0; j < w; ++j) {
            /* full[i+oh][j+ow] = sum_{u,v} image[i+oh-u][j+ow-v] * filter[u][v] */
            double acc = 0.0;
            for (int u = 0; u < fh; ++u) {
                int ii = i + oh - u;
                if (ii < 0 || ii >= h) continue;
                const double *irow = image + (int64_t)ii * w;
                const double *frow = filter + (int64_t)u * fw;
                for (int v = 0; v < fw; ++v) {
                    int jj = j + ow - v;
                    if (jj < 0 || jj >= w) continue;
                    acc += irow[jj] * frow[v];
                }
            }
            out[(int64_t)(i - row0) * w + j] = acc;
        }
}

/* ------------------------------------------------------------------ metrics */
double oracle_err_capped(const double *tested, const double *standard, int64_t n, int64_t batch)
{
    double error = 0.0;
    for (int64_t i = 0; i < batch; ++i)
        for (int64_t j = 0; j < n; ++j)
            for (int c = 0; c < 2; ++c) {
                double t = tested[c + 2 * j + 2 * n * i];
                double s = standard[c + 2 * j + 2 * n * i];
                double e = fabs((t - s) / s);
                error += (e < 1.0) ? e : 1.0; /* NaN (0/0) counts as 1, like std::min(1, NaN) */
            }
    return error / (double)n / (double)batch;
}

void oracle_err_l2max(const double *tested, const double *standard, int64_t count, double *rel_l2,
                      double *rel_max)
{
    double num = 0.0, den = 0.0, dmax = 0.0, smax = 0.0;
    for (int64_t i = 0; i < count; ++i) {
        double d = tested[i] - standard[i];
        num += d * d;
        den += standard[i] * standard[i];
        if (fabs(d) > dmax || d != d) dmax = (d != d) ? INFINITY : fabs(d);
        if (fabs(standard[i]) > smax) smax = fabs(standard[i]);
    }
    *rel_l2 = den > 0 ? sqrt(num / den) : sqrt(num);
    *rel_max = smax > 0 ? dmax / smax : dmax;
}
