/*
 * Drop-in for the reference's 2D matrix-core FFT interface (tcfft_half_2d.h:1-24).
 *
 *   struct tcfftHandle_2d   tcfft_half_2d.h:7-18   same public fields + opaque `impl`
 *   tcfftCreate_2d          tcfft_half_2d.cu:1151  fp16 MFMA plan, [batch][nx][ny] layout,
 *                                                  input digit-reversed along x and y
 *                                                  (rev_x, rev_y as in setup_2d :1264-1271)
 *   tcfftExec_2d            tcfft_half_2d.cu:1128  in place on device half2 data
 *   setup_2d/doit_2d/finalize_2d  tcfft_half_2d.cu:1253-1291 (natural-order FP64 host data)
 *
 * Unlike the reference (which only runs 256/512 x 256/512/1024, tcfft_half_2d.cu:1164-1182),
 * any power-of-two side from 16 to 4096 is accepted.
 */
#ifndef TCFFT_HALF_2D_H_MI355X
#define TCFFT_HALF_2D_H_MI355X

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tcfftHandle_2d {
    int Nx, Ny, N_batch;
    int radices_x[3];
    int radices_y[3];
    int n_radices_x, n_radices_y;
    int mergings[2];
    void *impl;
} tcfftHandle_2d;

void tcfftCreate_2d(tcfftHandle_2d *plan, int nx, int ny, int n_batch);
void tcfftExec_2d(tcfftHandle_2d plan, void *data);
void tcfftDestroy_2d(tcfftHandle_2d *plan); /* no counterpart in the reference */

void setup_2d(double *data, int nx, int ny, int n_batch);
void doit_2d(int iter);
void finalize_2d(double *result, int N_batch);

#ifdef __cplusplus
}
#endif
#endif
