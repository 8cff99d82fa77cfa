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
 *                                                     over num_streams streams
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
    ~fft_conv2d_half() { release(); }
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
        for (mfft_plan p : plans_) mfft_plan_destroy(p);
        plans_.clear();
    }

    void c2c(const double *input, double *out, int x, int y, int batch, int num_streams, int dir)
    {
        if (num_streams < 1) num_streams = 1;
        if (num_streams > batch) num_streams = batch;
        const int per = (batch + num_streams - 1) / num_streams;  // the reference drops batch % num_streams
        mfft_desc d;
        mfft_desc_default(&d);
        d.precision = MFFT_PREC_SPLIT_FP16;
        d.scale = MFFT_SCALE_NONE;
        d.direction = dir;
        d.in_fmt = d.out_fmt = MFFT_FMT_C32;
        mfft_plan plan = nullptr;
        check(mfft_plan_create_2d(&plan, y, x, per, &d), "fft_2d_c2c plan");
        const size_t elems = (size_t)x * y;
        std::vector<hipStream_t> st(num_streams);
        void *d64 = nullptr, *d32 = nullptr;
        check_hip(hipMalloc(&d64, elems * batch * 16));
        check_hip(hipMalloc(&d32, elems * (size_t)per * num_streams * 8));
        for (auto &s : st) check_hip(hipStreamCreate(&s));
        // one plan per stream: plans own their workspace
        std::vector<mfft_plan> sp(num_streams, nullptr);
        sp[0] = plan;
        for (int i = 1; i < num_streams; ++i) check(mfft_plan_create_2d(&sp[i], y, x, per, &d), "fft_2d_c2c plan");
        for (int i = 0; i < num_streams; ++i) {
            int b0 = i * per, nb = batch - b0 < per ? batch - b0 : per;
            if (nb <= 0) break;
            char *h_in = (char *)input + (size_t)b0 * elems * 16;
            char *g64 = (char *)d64 + (size_t)b0 * elems * 16;
            char *g32 = (char *)d32 + (size_t)i * per * elems * 8;
            check_hip(hipMemcpyAsync(g64, h_in, (size_t)nb * elems * 16, hipMemcpyHostToDevice, st[i]));
            check(mfft_stage_convert(g64, MFFT_FMT_C64, g32, MFFT_FMT_C32, elems, nb, nullptr, 0, nullptr, 0, 1.0, st[i]),
                  "stage");
            if (nb == per) {
                check(mfft_exec(sp[i], g32, g32, st[i]), "exec");
            } else {
                mfft_plan tail = nullptr;  // short last block
                check(mfft_plan_create_2d(&tail, y, x, nb, &d), "tail plan");
                check(mfft_exec(tail, g32, g32, st[i]), "exec");
                check_hip(hipStreamSynchronize(st[i]));
                mfft_plan_destroy(tail);
            }
            check(mfft_stage_convert(g32, MFFT_FMT_C32, g64, MFFT_FMT_C64, elems, nb, nullptr, 0, nullptr, 0, 1.0, st[i]),
                  "stage");
            check_hip(hipMemcpyAsync((char *)out + (size_t)b0 * elems * 16, g64, (size_t)nb * elems * 16,
                                     hipMemcpyDeviceToHost, st[i]));
        }
        for (auto &s : st) {
            check_hip(hipStreamSynchronize(s));
            check_hip(hipStreamDestroy(s));
        }
        for (mfft_plan p : sp) mfft_plan_destroy(p);
        check_hip(hipFree(d64));
        check_hip(hipFree(d32));
    }
};

#endif
