/*
 * Drop-in for the reference's 1D matrix-core FFT interface (tcfft_half.h:1-24) and its
 * comparator harness (cufft.h:1-6), on the MI355X library (include/mfft.h).
 *
 *   reference                                   here
 *   struct tcfftHandle      tcfft_half.h:7-18   same public fields; the kernel tables and
 *                                               device DFT matrices become `impl` (opaque)
 *   tcfftCreate             tcfft_half.cu:1235  mfft plan: fp16 MFMA, unnormalised, in place,
 *                                               digit-reversed input (gen_rev of `radices`)
 *   tcfftExec               tcfft_half.cu:1177  same contract: device half2 data in place,
 *                                               input digit-reversed, output natural order
 *   setup/doit/finalize     tcfft_half.cu:1376  same global harness; natural-order FP64 host
 *                                               data, staged to fp16 on the device
 *   setup_cu/doit_cu/...    cufft.cpp:12-89     the plain fp32 GPU FFT comparator of this
 *                                               library (no vendor FFT)
 *
 * Every function keeps the reference's void return; failures are reported through
 * tcfftGetLastError() (0 = ok, else an MFFT_ERR_* code) and mfft_last_error().
 * Data pointers are plain `void *` so both `half *` and `__half *` callers bind.
 */
#ifndef TCFFT_HALF_H_MI355X
#define TCFFT_HALF_H_MI355X

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tcfftHandle {
    int N, N_batch;
    int radices[9];
    int n_radices;
    int mergings[3];
    int n_mergings;
    void *impl;  /* opaque plan (replaces layer_0/layer_1/F_real/F_imag) */
} tcfftHandle;

void tcfftCreate(tcfftHandle *plan, int n, int n_batch);
void tcfftExec(tcfftHandle plan, void *data);
void tcfftDestroy(tcfftHandle *plan); /* no counterpart in the reference (it leaks plans) */
int tcfftGetLastError(void);

/* harness over process-global state (tcfft_half.cu:1347-1413): not thread-safe */
void setup(double *data, int n, int n_batch);
void doit(int iter);
void finalize(double *result, int N_batch);

/* plain fp32 GPU FFT comparator harness (cufft.h:1-6) */
void setup_cu(double *data, int n, int n_batch);
void doit_cu(int iter);
void finalize_cu(double *result, int n, int n_batch);
void setup_cu_2d(double *data, int nx, int ny, int n_batch);
void doit_cu_2d(int iter);
void finalize_cu_2d(double *result, int nx, int ny, int n_batch);

#ifdef __cplusplus
}
#endif
#endif
