/*
 * FP64 CPU oracle for the matrix-core FFT / FFT-conv2d path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or calls
 * this code; it is the checker used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.
 *
 * Every function restates behaviour of the reference (yuelingfan/fft_demo) and
 * cites the file:line it follows. Conventions:
 *   - complex data is interleaved (re, im) doubles, element index fastest, then batch
 *     (main.cpp:68-80);
 *   - forward transform uses W = exp(-2*pi*i/N), unnormalised (main.cpp:71, FFTW_FORWARD);
 *     sign = -1 is forward, sign = +1 is the unnormalised inverse
 *     (tensorcore_fft_conv2d_kernel.cu:47-82);
 *   - 2D layout is [batch][x][y] with y contiguous (main.cpp:82-94,
 *     tcfft_half_2d.cu:1264-1271).
 */
#ifndef MFFT_ORACLE_H
#define MFFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mixed-radix digit-reversal permutation; restates gen_rev, tcfft_half.cu:1351-1374.
 * rev[i] is the natural-order index that must be stored at position i. Returns 0, or
 * -1 if the radices do not multiply to n. */
int oracle_gen_rev(int n, const int *radices, int n_radices, int *rev);

/* The raw SplitMix64 stream (Steele, Lea and Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014; Vigna's reference splitmix64.c): count outputs from `seed`. */
void oracle_splitmix64(uint64_t seed, uint64_t *out, int64_t count);

/* SplitMix64 uniform[-1,1] generator (seeded stream). x = splitmix64(state);
 * u = (x >> 11) * 2^-53; value = lo + (hi - lo) * u. */
void oracle_fill_uniform(double *dst, int64_t count, uint64_t seed, double lo, double hi);

/* Round every value to the nearest IEEE binary16 (ties to even), keeping it as double. */
void oracle_round_fp16(double *data, int64_t count);
/* Round every value to the nearest bfloat16 (ties to even), keeping it as double. */
void oracle_round_bf16(double *data, int64_t count);

/* Brute-force O(n^2) DFT, exact integer phase reduction (jk mod n). sign -1 forward. */
void oracle_dft(const double *in, double *out, int64_t n, int64_t batch, int sign);

/* Mixed-radix Stockham FFT in FP64 (any n >= 1; radices 4,2,3,5,7 and generic primes).
 * threads <= 0 uses all OpenMP threads. Returns 0 on success. */
int oracle_fft(const double *in, double *out, int64_t n, int64_t batch, int sign, int threads);

/* 2D FFT on [batch][nx][ny] (y contiguous): rows along y, then columns along x. */
int oracle_fft2d(const double *in, double *out, int64_t nx, int64_t ny, int64_t batch, int sign,
                 int threads);

/* Shape of the reference's CPU timing harness (main.cpp:68-80): one plan, a single-thread
 * batch loop that memcpy's each transform into a work buffer, transforms it in place
 * and copies it out. */
int oracle_fft_batch_loop(const double *data, double *result, int64_t n, int64_t batch);

/* 'valid' cross-correlation of one image with one filter; restates
 * naive_fft_conv2d_ss_sf_valid, main.cpp:97-116. out is (sy-fy+1) x (sx-fx+1). */
void oracle_corr2d_valid(const double *signal, const double *filter, double *out, int sy, int sx,
                         int fy, int fx);

/* 'same' 2D convolution (true convolution, filter flipped), output h x w, centred on the
 * full convolution with offset ((fh-1)/2, (fw-1)/2), zero boundary. Rows [row0, row1)
 * only (pass 0, h for the whole image). */
/* Direct FP64 'full' 2D convolution out[Y][X] = sum_{i,j} a[Y-i][X-j] * b[i][j],
 * out of size (ah+bh-1) x (aw+bw-1). The input gradient of the 'valid' correlation layer
 * (the quantity tensorcore_fft_conv2d.cpp:400-442 computes for conv2d_backward). */
void oracle_conv2d_full(const double *a, const double *b, double *out, int ah, int aw, int bh, int bw);

void oracle_conv2d_same(const double *image, const double *filter, double *out, int h, int w,
                        int fh, int fw, int row0, int row1, int threads);

/* Capped relative error of the reference, main.cpp:118-134: sum over re and im of
 * min(1, |(t - s) / s|), divided by n * batch. */
double oracle_err_capped(const double *tested, const double *standard, int64_t n, int64_t batch);

/* rel_l2 = ||t - s||_2 / ||s||_2 ; rel_max = max|t - s| / max|s| over `count` doubles. */
void oracle_err_l2max(const double *tested, const double *standard, int64_t count, double *rel_l2,
                      double *rel_max);

#ifdef __cplusplus
}
#endif
#endif
