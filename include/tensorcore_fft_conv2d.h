/*
 * Drop-in for the reference's FFT-convolution operator (tensorcore_fft_conv2d.h:1-197):
 * the same names and argument meaning, implemented header-only over the C ABI of
 * include/mfft.h (mfft_conv2d_* and the 2D plans). C++ only.
 *
 *   reference                                          here
 *   enum __GPU_ARCH__ {__VOLTA__, __AMPERE__}  :12-16  kept for source compatibility, ignored
 *   struct fftPlan                              :26-40  2D plan over mfft_plan (fft = 1 / -1)
 *   TensorCoreFFTCreate / TensorCoreFFTExec  _kernel.cu:1495-1648
 *                                                     mfft 2D plan, fp16 matrix-core,
 *                                                     unnormalised, digit-reversed input
 *   fft_conv2d_half(signal, filter, S, F, sy, sx, fy, fx, GPU)  .cpp:12-58
 *                                                     host FP64 inputs, 'valid' correlation
 *                                                     of S channels with F x S filters
 *   conv2d_forward(double *out)                  .cpp:283-290   host FP64 output
 *   conv2d_forward(in, w, out, S, F, sy, sx, fy, fx, GPU)  .cpp:291-363
 *   fft_2d_c2c / ifft_2d_c2c                     .cpp:79-214   host FP64 complex batches
 *                                                     over num_streams streams, staged
 *                                                     through pinned host memory
 *   conv2d_backward(double *delta_in)             .cpp:364-442  host FP64 dL/dout in;
 *                                                     delta_for_filter [F][S][fy][fx] and
 *                                                     delta_for_signal [S][sy][sx] out
 * Differences by design: any sizes (the reference pads only to 256 or 512), fp32 spectra
 * from split-fp16 transforms instead of fp16 spectra with power-of-10 quantisation, no
 * double free of the device buffers (the reference frees them in fft_pointwise and again
 * in the destructor, _kernel.cu:1110-1111 / .cpp:73-74), and the backward pass computes
 * the gradients of the 'valid' correlation exactly (input gradient = 'full' convolution
 * of dL/dout with the filters, filter gradient = 'valid' correlation of the signal with
 * dL/dout) on the forward's padded size, with the output arrays owned by the object.
 */
#ifndef TENSORCORE_FFT_CONV2D_H_MI355X
#define TENSORCORE_FFT_CONV2D_H_MI355X

#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mfft.h"

enum __GPU_ARCH__ { __VOLTA__, __AMPERE__ };
enum __OBJ__ { __SIGNAL__, __FILTER__ };

/* fft = 1 -> forward, fft = -1 -> inverse (unnormalised), as in the reference */
struct fftPlan {
    __GPU_ARCH__ GPU = __AMPERE__;
    __OBJ__ OBJ = __SIGNAL__;
    int fft = 1;
    int Nx = 0, Ny = 0, N_batch = 0;
    mfft_plan impl = nullptr;
};

class fft_conv2d_half {
  public:
    __GPU_ARCH__ GPU = __AMPERE__;
    int sx = 0, sy = 0, fx = 0, fy = 0, S = 0, F = 0;
    int padded_x = 0, padded_y = 0, output_x = 0, output_y = 0;
    double *signal = nullptr, *filter = nullptr, *output = nullptr;

    fft_conv2d_half() {}
    fft_conv2d_half(int in_channel, int out_channel, int kernel_size, __GPU_ARCH__ gpu)
        : GPU(gpu), fx(kernel_size), fy(kernel_size), S(in_channel), F(out_channel)
    {
    }
    fft_conv2d_half(double *signal_, double *filter_, int S_, int F_, int sy_, int sx_, int fy_, int fx_,
                    __GPU_ARCH__ gpu)
    {
        init(signal_, filter_, S_, F_, sy_, sx_, fy_, fx_, gpu);
    }
    ~fft_conv2d_half()
    {
        release();
        for (mfft_plan p : plans_) mfft_plan_destroy(p);
        c2c_release();
    }
    fft_conv2d_half(const fft_conv2d_half &) = delete;
    fft_conv2d_half &operator=(const fft_conv2d_half &) = delete;

    /* 'valid' cross-correlation of the constructor's signal with its filters: out holds F
     * maps of output_y x output_x doubles (host memory). */
    void conv2d_forward(double *out)
    {
        output = out;
        check(mfft_conv2d_forward(op_, d_signal_, MFFT_REAL_F64, d_out_, MFFT_REAL_F64, nullptr), "forward");
        size_t n = (size_t)F * output_x * output_y;
        check_hip(hipMemcpy(out, d_out_, n * sizeof(double), hipMemcpyDeviceToHost));
    }

    void conv2d_forward(double *input, double *weight, double *out, int S_, int F_, int sy_, int sx_, int fy_, int fx_,
                        __GPU_ARCH__ gpu)
    {
        release();
        init(input, weight, S_, F_, sy_, sx_, fy_, fx_, gpu);
        conv2d_forward(out);
    }

    /* dL/dout (F maps of output_y x output_x doubles, host) -> delta_for_filter and
     * delta_for_signal (host arrays owned by this object) */
    double *delta_for_filter = nullptr, *delta_for_signal = nullptr;
    void conv2d_backward(double *delta_in)
    {
        const size_t nd = (size_t)F * output_x * output_y, ng = (size_t)F * S * fx * fy, ns = (size_t)S * sx * sy;
        void *d_delta = nullptr, *d_gw = nullptr, *d_gi = nullptr;
        check_hip(hipMalloc(&d_delta, nd * sizeof(double)));
        check_hip(hipMalloc(&d_gw, ng * sizeof(double)));
        check_hip(hipMalloc(&d_gi, ns * sizeof(double)));
        check_hip(hipMemcpy(d_delta, delta_in, nd * sizeof(double), hipMemcpyHostToDevice));
        check(mfft_conv2d_backward(op_, d_signal_, d_delta, MFFT_REAL_F64, d_gi, d_gw, nullptr), "conv2d_backward");
        grad_w_.resize(ng);
        grad_in_.resize(ns);
        check_hip(hipMemcpy(grad_w_.data(), d_gw, ng * sizeof(double), hipMemcpyDeviceToHost));
        check_hip(hipMemcpy(grad_in_.data(), d_gi, ns * sizeof(double), hipMemcpyDeviceToHost));
        (void)hipFree(d_delta);
        (void)hipFree(d_gw);
        (void)hipFree(d_gi);
        delta_for_filter = grad_w_.data();
        delta_for_signal = grad_in_.data();
    }

    void TensorCoreFFTCreate(fftPlan *plan, int nx, int ny, int n_batch, int fft, __GPU_ARCH__ gpu)
    {
        mfft_desc d;
        mfft_desc_default(&d);
        d.precision = MFFT_PREC_FP16;
        d.scale = MFFT_SCALE_NONE;
        d.direction = fft == 1 ? MFFT_FORWARD : MFFT_INVERSE;
        d.in_order = MFFT_ORDER_DIGIT_REVERSED;
        plan->GPU = gpu;
        plan->fft = fft;
        plan->Nx = nx;
        plan->Ny = ny;
        plan->N_batch = n_batch;
        check(mfft_plan_create_2d(&plan->impl, nx, ny, n_batch, &d), "TensorCoreFFTCreate");
        plans_.push_back(plan->impl);
    }
    void TensorCoreFFTExec(fftPlan plan, void *data, hipStream_t stream)
    {
        check(mfft_exec(plan.impl, data, data, (void *)stream), "TensorCoreFFTExec");
    }
    void TensorCoreFFTExec(fftPlan plan, void *data) { TensorCoreFFTExec(plan, data, nullptr); }

    /* batched 2D C2C on host FP64 interleaved data ([batch][y][x], x contiguous as in the
     * reference's pitched copies), split over num_streams HIP streams (H2D, convert, FFT,
     * convert, D2H overlap across streams); output unnormalised (forward) */
    void fft_2d_c2c(double *input, double *out, int x, int y, int batch, __GPU_ARCH__ gpu, int num_streams = 8,
                    bool benchmark = false)
    {
        (void)gpu;
        (void)benchmark;
        c2c(input, out, x, y, batch, num_streams, MFFT_FORWARD);
    }
    void ifft_2d_c2c(double *input, double *out, int x, int y, int batch, __GPU_ARCH__ gpu, int num_streams = 8,
                     bool benchmark = false)
    {
        (void)gpu;
        (void)benchmark;
        c2c(input, out, x, y, batch, num_streams, MFFT_INVERSE);
    }

  private:
    mfft_conv2d op_ = nullptr;
    void *d_signal_ = nullptr, *d_out_ = nullptr;
    std::vector<double> grad_w_, grad_in_;
    std::vector<mfft_plan> plans_;

    static void check(int rc, const char *what)
    {
        if (rc != MFFT_OK) throw std::runtime_error(std::string(what) + ": " + mfft_last_error());
    }
    static void check_hip(hipError_t e)
    {
        if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
    }

    void init(double *signal_, double *filter_, int S_, int F_, int sy_, int sx_, int fy_, int fx_, __GPU_ARCH__ gpu)
    {
        signal = signal_;
        filter = filter_;
        S = S_;
        F = F_;
        sy = sy_;
        sx = sx_;
        fy = fy_;
        fx = fx_;
        GPU = gpu;
        output_x = sx - fx + 1;
        output_y = sy - fy + 1;
        mfft_conv2d_desc d;
        mfft_conv2d_desc_default(&d);
        d.batch = 1;
        d.channels = S;
        d.filters = F;
        d.h = sy;
        d.w = sx;
        d.fh = fy;
        d.fw = fx;
        d.mode = MFFT_CONV_VALID_CORR;
        check(mfft_conv2d_create(&op_, &d), "fft_conv2d_half");
        mfft_conv2d_info inf;
        mfft_conv2d_info_get(op_, &inf);
        padded_y = inf.pad_h;
        padded_x = inf.pad_w;
        void *d_filter = nullptr;
        size_t ns = (size_t)S * sx * sy, nf = (size_t)F * S * fx * fy, no = (size_t)F * output_x * output_y;
        check_hip(hipMalloc(&d_signal_, ns * sizeof(double)));
        check_hip(hipMalloc(&d_out_, no * sizeof(double)));
        check_hip(hipMalloc(&d_filter, nf * sizeof(double)));
        check_hip(hipMemcpy(d_signal_, signal, ns * sizeof(double), hipMemcpyHostToDevice));
        check_hip(hipMemcpy(d_filter, filter, nf * sizeof(double), hipMemcpyHostToDevice));
        check(mfft_conv2d_set_filters(op_, d_filter, MFFT_REAL_F64, nullptr), "set filters");
        check_hip(hipDeviceSynchronize());
        check_hip(hipFree(d_filter));
    }

    void release()
    {
        if (op_) mfft_conv2d_destroy(op_);
        op_ = nullptr;
        if (d_signal_) (void)hipFree(d_signal_);
        if (d_out_) (void)hipFree(d_out_);
        d_signal_ = d_out_ = nullptr;
    }

    // fft_2d_c2c state, kept across calls while the shape, split and direction stay the same:
    // one plan for a stream's block of transforms (external workspace), per-stream device
    // buffers and workspaces, and pinned host staging (the caller's arrays are pageable, and
    // a copy from pageable memory does not overlap with other streams' work)
    struct C2C {
        int x = 0, y = 0, per = 0, nst = 0, dir = 0;
        mfft_plan plan = nullptr;
        std::vector<hipStream_t> st;
        std::vector<void *> dev64, dev32, ws, pin_in, pin_out;
    } c2c_;

    void c2c_release()
    {
        C2C &c = c2c_;
        for (hipStream_t s : c.st) (void)hipStreamSynchronize(s);
        if (c.plan) mfft_plan_destroy(c.plan);
        for (auto *v : {&c.dev64, &c.dev32, &c.ws})
            for (void *p : *v) (void)hipFree(p);
        for (auto *v : {&c.pin_in, &c.pin_out})
            for (void *p : *v) (void)hipHostFree(p);
        for (hipStream_t s : c.st) (void)hipStreamDestroy(s);
        c2c_ = C2C();
    }

    void c2c_setup(int x, int y, int per, int nst, int dir)
    {
        C2C &c = c2c_;
        if (c.plan && c.x == x && c.y == y && c.per == per && c.nst == nst && c.dir == dir) return;
        c2c_release();
        mfft_desc d;
        mfft_desc_default(&d);
        d.precision = MFFT_PREC_SPLIT_FP16;
        d.scale = MFFT_SCALE_NONE;
        d.direction = dir;
        d.in_fmt = d.out_fmt = MFFT_FMT_C32;
        d.external_workspace = 1;
        check(mfft_plan_create_2d(&c.plan, y, x, per, &d), "fft_2d_c2c plan");
        mfft_info inf;
        check(mfft_plan_info(c.plan, &inf), "fft_2d_c2c plan info");
        c.x = x;
        c.y = y;
        c.per = per;
        c.nst = nst;
        c.dir = dir;
        const size_t block = (size_t)x * y * per;
        c.st.resize(nst);
        c.dev64.assign(nst, nullptr);
        c.dev32.assign(nst, nullptr);
        c.ws.assign(nst, nullptr);
        c.pin_in.assign(nst, nullptr);
        c.pin_out.assign(nst, nullptr);
        for (int i = 0; i < nst; ++i) {
            check_hip(hipStreamCreateWithFlags(&c.st[i], hipStreamNonBlocking));
            check_hip(hipMalloc(&c.dev64[i], block * 16));
            check_hip(hipMalloc(&c.dev32[i], block * 8));
            check_hip(hipMalloc(&c.ws[i], inf.workspace_bytes > 0 ? (size_t)inf.workspace_bytes : 16));
            check_hip(hipHostMalloc(&c.pin_in[i], block * 16, hipHostMallocDefault));
            check_hip(hipHostMalloc(&c.pin_out[i], block * 16, hipHostMallocDefault));
        }
    }

    // Batch split into num_streams blocks, one per stream. Block i: host copy of the
    // caller's input into pinned memory, then on stream i the copy to the device, fp64 ->
    // fp32 on the device, the transform (a share of the one plan, the stream's own
    // workspace), fp32 -> fp64, the copy back to pinned memory; the host copies block i's
    // result out while later blocks run. Unlike the reference (which drops batch %
    // num_streams transforms, tensorcore_fft_conv2d.cpp:84) every transform is done.
    void c2c(const double *input, double *out, int x, int y, int batch, int num_streams, int dir)
    {
        if (batch < 1 || x < 2 || y < 2) throw std::runtime_error("fft_2d_c2c: bad sizes");
        if (num_streams < 1) num_streams = 1;
        if (num_streams > batch) num_streams = batch;
        const int per = (batch + num_streams - 1) / num_streams;
        c2c_setup(x, y, per, num_streams, dir);
        C2C &c = c2c_;
        const size_t elems = (size_t)x * y;
        const int blocks = (batch + per - 1) / per;
        for (int i = 0; i < blocks; ++i) {
            const int b0 = i * per, nb = batch - b0 < per ? batch - b0 : per;
            const size_t bytes = (size_t)nb * elems * 16;
            memcpy(c.pin_in[i], (const char *)input + (size_t)b0 * elems * 16, bytes);
            check_hip(hipMemcpyAsync(c.dev64[i], c.pin_in[i], bytes, hipMemcpyHostToDevice, c.st[i]));
            check(mfft_stage_convert(c.dev64[i], MFFT_FMT_C64, c.dev32[i], MFFT_FMT_C32, elems, nb, nullptr, 0, nullptr,
                                     0, 1.0, c.st[i]),
                  "fft_2d_c2c stage in");
            check(mfft_exec_ex(c.plan, c.dev32[i], c.dev32[i], nb, c.ws[i], c.st[i]), "fft_2d_c2c exec");
            check(mfft_stage_convert(c.dev32[i], MFFT_FMT_C32, c.dev64[i], MFFT_FMT_C64, elems, nb, nullptr, 0, nullptr,
                                     0, 1.0, c.st[i]),
                  "fft_2d_c2c stage out");
            check_hip(hipMemcpyAsync(c.pin_out[i], c.dev64[i], bytes, hipMemcpyDeviceToHost, c.st[i]));
        }
        for (int i = 0; i < blocks; ++i) {
            const int b0 = i * per, nb = batch - b0 < per ? batch - b0 : per;
            check_hip(hipStreamSynchronize(c.st[i]));
            memcpy((char *)out + (size_t)b0 * elems * 16, c.pin_out[i], (size_t)nb * elems * 16);
        }
    }
};

#endif
