// FFT-based 2D convolution operator (C ABI in include/mfft.h).
//
// Reference operator: class fft_conv2d_half (tensorcore_fft_conv2d.h:41-197); forward
// path = ctor (tensorcore_fft_conv2d.cpp:12-58: pad, digit-reverse, FFT signal and
// filter), TENSORCORE_FFT_CONV_2D (:250-273: pointwise, inverse FFT, dequantise) and
// readout_fwd (:275-282: crop at fx-1, scale 1/(Nx*Ny)). Differences by design: any image
// size (padding to 2^a*{1,3,5,9,15,17} instead of 256/512 only, :23-32), 'same' convolution
// besides the reference's 'valid' correlation, batches with pairs of real images sharing
// a complex transform, fp32 spectra with per-stage-scaled split-fp16 transforms instead of
// fp16 spectra with per-filter power-of-10 quantisation, resident filter spectra reused
// across calls, and no per-call allocation.
#include "../../include/mfft.h"
#include "mfft_internal.h"

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

namespace mfft {
hipError_t conv_pack_pairs(const void *src, int f64, void *dst, int fmt, int B, int C, int h, int w, int Ph, int Pw,
                           hipStream_t s);
hipError_t conv_pack_filters(const void *src, int f64, float2 *dst, int FC, int fh, int fw, int Ph, int Pw, int rotate,
                             hipStream_t s);
hipError_t conv_pointwise(const void *X, const float2 *H, void *Y, int fmt, long long pairs, int C, int F, long long P,
                          hipStream_t s, int yfmt = -1);
hipError_t conv_quantise(const float2 *R, void *Q, float *partial, float *scale, long long maps, long long P,
                         int nchunk, hipStream_t s);
hipError_t conv_crop(const void *Y, int fmt, void *out, int f64, int B, int F, int oh, int ow, int Ph, int Pw, int oy,
                     int ox, float scale, hipStream_t s, const float *map_scale = nullptr);
hipError_t conv_pointwise_gw(const float2 *X, const float2 *D, float2 *G, int pairs, int C, int F, int Ph, int Pw,
                             hipStream_t s);
hipError_t conv_pointwise_gi(const float2 *D, const float2 *H, float2 *R, long long pairs, int C, int F, long long P,
                             hipStream_t s);
hipError_t conv_crop_cyclic(const float2 *Y, void *out, int f64, int B, int C, int oh, int ow, int Ph, int Pw, int oy,
                            int ox, float scale, hipStream_t s);
hipError_t conv_crop_real(const float2 *Y, void *out, int f64, long long planes, int oh, int ow, int Ph, int Pw,
                          float scale, hipStream_t s);
hipError_t conv_pack_two(const void *a, const void *b, int f64, void *dst, int fmt, int h, int w, int Ph, int Pw,
                         hipStream_t s);
hipError_t conv_pointwise_two(const void *X, const float2 *H0, const float2 *H1, void *Y, int fmt, int Ph, int Pw,
                              hipStream_t s);
hipError_t conv_crop_two(const void *Y, int fmt, void *out0, void *out1, int f64, int oh, int ow, int Pw, int oy,
                         int ox, hipStream_t s);
}  // namespace mfft

struct mfft_conv2d_s {
    mfft_conv2d_desc d{};
    int Ph = 0, Pw = 0, oh = 0, ow = 0, oy = 0, ox = 0;
    long long pairs = 0;
    int spec_fmt = MFFT_FMT_C32;  // image and product spectra X, Y (the filter spectra H stay fp32)
    // fp32 spectra, fp16 pass boundaries (MFFT_CONV_SPEC_FP32_MID_FP16): the image transforms
    // run out of place, so a transform whose input is not fused reads Y and writes X, and the
    // inverse of a pointwise product in Y writes X
    bool mid16 = false;
    int res_fmt = MFFT_FMT_C32;  // the inverse transform's output, which the crop reads
    // the reference's method (MFFT_CONV_SPEC_REFERENCE): unscaled fp16 transforms, the fp32
    // product spectra in Y quantised per output map into Q (fp16) with the scales in qscale
    bool refq = false;
    void *Q = nullptr;
    float *qscale = nullptr, *qpart = nullptr;
    int qchunks = 1;
    mfft_plan fwd_img = nullptr, fwd_filt = nullptr, inv = nullptr;
    float2 *H = nullptr, *X = nullptr, *Y = nullptr;
    // backward (built on its first call): delta transform, inverses over (filter, channel)
    // and (pair, channel) batches, and the filter-gradient spectra
    mfft_plan fwd_delta = nullptr, inv_fc = nullptr, inv_pc = nullptr;
    float2 *G = nullptr;
    bool have_filters = false;
    long long bytes = 0, spec_bytes = 0;
    std::mutex mu;
    mfft::StreamOrder order;  // streams that used H, X, Y, G
};

namespace {

// Errors of the conv entry points go to the thread's mfft_last_error() message.
int cfail(int code, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    mfft::set_last_error(buf);
    return code;
}

// Padded sizes odd * 2^a whose odd part has a streaming pass of its own (the planner then
// runs the power of two on the matrix-core kernels); products such as 51 = 3 * 17 would
// fall back to the mixed-radix tile kernel.
bool good(int n)
{
    if (n < 2 || n > 4096 * 256) return false;
    while (n % 2 == 0) n /= 2;
    return n == 1 || n == 3 || n == 5 || n == 9 || n == 15 || n == 17;
}

// bytes of one image or product spectrum element
long long spec_eb(const mfft_conv2d_s *op) { return op->spec_fmt == MFFT_FMT_C16 ? 4 : 8; }

// One call on the operator: holds its mutex and orders its buffers (H, X, Y, G) after the
// previous call's when that ran on another stream.
struct OpCall {
    mfft_conv2d op;
    hipStream_t s;
    std::lock_guard<std::mutex> lk;
    hipError_t err;
    OpCall(mfft_conv2d o, hipStream_t st) : op(o), s(st), lk(o->mu), err(o->order.acquire(st)) {}
    ~OpCall() { (void)op->order.release(s); }
};

int plan_info_bytes(mfft_plan p)
{
    mfft_info inf;
    if (mfft_plan_info(p, &inf)) return 0;
    return (int)(inf.workspace_bytes >> 10);  // KiB
}

}  // namespace

extern "C" {

int mfft_good_size(int n)
{
    for (int m = n < 2 ? 2 : n; m <= 4096 * 256; ++m)
        if (good(m)) {
            // keep sizes the transform handles in at most two passes per axis
            return m;
        }
    return -1;
}

int mfft_conv2d_desc_default(mfft_conv2d_desc *d)
{
    if (!d) return MFFT_ERR_INVALID;
    memset(d, 0, sizeof *d);
    d->batch = d->channels = d->filters = 1;
    d->mode = MFFT_CONV_SAME;
    d->precision = MFFT_PREC_SPLIT_FP16;
    return MFFT_OK;
}

int mfft_conv2d_destroy(mfft_conv2d op)
{
    if (!op) return MFFT_OK;
    (void)hipDeviceSynchronize();
    if (op->fwd_img) mfft_plan_destroy(op->fwd_img);
    if (op->fwd_filt) mfft_plan_destroy(op->fwd_filt);
    if (op->inv) mfft_plan_destroy(op->inv);
    for (mfft_plan p : {op->fwd_delta, op->inv_fc, op->inv_pc})
        if (p) mfft_plan_destroy(p);
    (void)hipFree(op->G);
    (void)hipFree(op->H);
    (void)hipFree(op->X);
    (void)hipFree(op->Y);
    (void)hipFree(op->Q);
    (void)hipFree(op->qscale);
    (void)hipFree(op->qpart);
    op->order.destroy();
    delete op;
    return MFFT_OK;
}

int mfft_conv2d_create(mfft_conv2d *out, const mfft_conv2d_desc *d)
{
    if (!out || !d) return cfail(MFFT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (d->batch < 1 || d->channels < 1 || d->filters < 1 || d->h < 1 || d->w < 1 || d->fh < 1 || d->fw < 1)
        return cfail(MFFT_ERR_INVALID, "bad sizes");
    if (d->mode != MFFT_CONV_SAME && d->mode != MFFT_CONV_VALID_CORR) return cfail(MFFT_ERR_INVALID, "mode");
    if (d->mode == MFFT_CONV_VALID_CORR && (d->fh > d->h || d->fw > d->w))
        return cfail(MFFT_ERR_INVALID, "filter larger than the image in 'valid' mode");
    if (d->spec_fp16 < MFFT_CONV_SPEC_FP32 || d->spec_fp16 > MFFT_CONV_SPEC_REFERENCE)
        return cfail(MFFT_ERR_INVALID, "spec_fp16");
    if (d->spec_fp16 == MFFT_CONV_SPEC_FP32_MID_FP16 && d->channels != 1)
        return cfail(MFFT_ERR_UNSUPPORTED, "fp16 pass boundaries need a single-channel operator");
    mfft_conv2d op = new mfft_conv2d_s;
    op->d = *d;
    op->refq = d->spec_fp16 == MFFT_CONV_SPEC_REFERENCE;
    op->spec_fmt = d->spec_fp16 == MFFT_CONV_SPEC_FP16 || op->refq ? MFFT_FMT_C16 : MFFT_FMT_C32;
    op->mid16 = d->spec_fp16 == MFFT_CONV_SPEC_FP32_MID_FP16;
    int need_h = d->h + d->fh - 1, need_w = d->w + d->fw - 1;
    if (d->mode == MFFT_CONV_VALID_CORR) {
        // circular wrap-around only touches outputs outside the 'valid' window
        need_h = d->h;
        need_w = d->w;
    }
    op->Ph = d->pad_h > 0 ? d->pad_h : mfft_good_size(need_h);
    op->Pw = d->pad_w > 0 ? d->pad_w : mfft_good_size(need_w);
    if (op->Ph < need_h || op->Pw < need_w) {
        mfft_conv2d_destroy(op);
        return cfail(MFFT_ERR_INVALID, "padding %dx%d smaller than %dx%d", op->Ph, op->Pw, need_h, need_w);
    }
    if (d->mode == MFFT_CONV_SAME) {
        op->oh = d->h;
        op->ow = d->w;
        op->oy = (d->fh - 1) / 2;
        op->ox = (d->fw - 1) / 2;
    } else {
        op->oh = d->h - d->fh + 1;
        op->ow = d->w - d->fw + 1;
        op->oy = d->fh - 1;
        op->ox = d->fw - 1;
    }
    op->pairs = (d->batch + 1) / 2;
    mfft_desc fd;
    mfft_desc_default(&fd);
    fd.precision = d->precision;
    fd.in_fmt = fd.out_fmt = op->spec_fmt;
    if (op->spec_fmt == MFFT_FMT_C16 || op->mid16) fd.mid_fmt = MFFT_FMT_C16;
    fd.workspace_limit = d->workspace_limit;
    fd.scale = MFFT_SCALE_STAGE;  // image spectra carry 1/(Ph*Pw): the inverse is then unit
    if (op->refq) {
        // the reference's transforms: fp16 operands, no scaling in either direction
        // (tensorcore_fft_conv2d_kernel.cu:109-1031); 1/(Ph*Pw) and the quantisation scale
        // are applied by the crop
        fd.precision = MFFT_PREC_FP16;
        fd.scale = MFFT_SCALE_NONE;
    }
    // fp16 pass boundaries with fp32 ends: the intermediates the passes keep in the output
    // buffer are fp16 as well (the image transforms then never run in place)
    auto create = [&](mfft_plan *pl, int64_t b) {
        return op->mid16 ? mfft::plan_create_2d_out_mid(pl, op->Ph, op->Pw, b, &fd)
                         : mfft_plan_create_2d(pl, op->Ph, op->Pw, b, &fd);
    };
    int rc = create(&op->fwd_img, op->pairs * d->channels);
    if (!rc) {
        fd.direction = MFFT_INVERSE;
        fd.scale = MFFT_SCALE_NONE;
        // fp16 pass boundaries: the inverse's output, read only by the crop, is fp16 words as
        // well (the spectra it reads stay fp32)
        if (op->mid16) fd.out_fmt = MFFT_FMT_C16;
        rc = create(&op->inv, op->pairs * d->filters);
        fd.out_fmt = op->spec_fmt;
    }
    op->res_fmt = op->mid16 ? MFFT_FMT_C16 : op->spec_fmt;
    if (!rc) {
        // the resident filter spectra stay fp32 whatever the image spectra
        fd.direction = MFFT_FORWARD;
        fd.in_fmt = fd.out_fmt = MFFT_FMT_C32;
        fd.mid_fmt = -1;
        rc = mfft_plan_create_2d(&op->fwd_filt, op->Ph, op->Pw, (int64_t)d->filters * d->channels, &fd);
    }
    if (rc) {
        std::string m = mfft_last_error();
        mfft_conv2d_destroy(op);
        return cfail(rc, "%s", m.c_str());
    }
    long long P = (long long)op->Ph * op->Pw;
    op->spec_bytes = (long long)d->filters * d->channels * P * 8;
    const long long eb = op->spec_fmt == MFFT_FMT_C16 ? 4 : 8;
    const long long maps = op->pairs * d->filters;
    long long xb = op->pairs * d->channels * P * eb, yb = maps * P * (op->refq ? 8 : eb);
    if (hipMalloc(&op->H, op->spec_bytes) != hipSuccess || hipMalloc(&op->X, xb) != hipSuccess ||
        hipMalloc(&op->Y, yb) != hipSuccess) {
        mfft_conv2d_destroy(op);
        return cfail(MFFT_ERR_NOMEM, "conv buffers (%lld + %lld + %lld bytes)", op->spec_bytes, xb, yb);
    }
    op->bytes = op->spec_bytes + xb + yb;
    if (op->refq) {
        // quantised product spectra, one scale per map, chunk maxima (one chunk per 16K values)
        op->qchunks = (int)std::max(1LL, std::min(256LL, P / 16384));
        const long long qb = maps * P * 4, sb = maps * 4, pb = maps * op->qchunks * 4;
        if (hipMalloc(&op->Q, qb) != hipSuccess || hipMalloc((void **)&op->qscale, sb) != hipSuccess ||
            hipMalloc((void **)&op->qpart, pb) != hipSuccess) {
            mfft_conv2d_destroy(op);
            return cfail(MFFT_ERR_NOMEM, "quantisation buffers (%lld bytes)", qb + sb + pb);
        }
        op->bytes += qb + sb + pb;
    }
    mfft_info inf;
    for (mfft_plan p : {op->fwd_img, op->fwd_filt, op->inv})
        if (!mfft_plan_info(p, &inf)) op->bytes += inf.workspace_bytes;
    *out = op;
    return MFFT_OK;
}

int mfft_conv2d_set_filters(mfft_conv2d op, const void *filters, int real_fmt, void *stream)
{
    if (!op || !filters) return cfail(MFFT_ERR_INVALID, "null argument");
    hipStream_t s = (hipStream_t)stream;
    OpCall call(op, s);
    if (call.err != hipSuccess) return cfail(MFFT_ERR_HIP, "stream order: %s", hipGetErrorString(call.err));
    const mfft_conv2d_desc &d = op->d;
    mfft::trace_launch(s, "conv_pack_filters", (long long)d.filters * d.channels * d.fh * d.fw * (real_fmt == MFFT_REAL_F64 ? 8 : 4),
                       op->spec_bytes);
    if (mfft::conv_pack_filters(filters, real_fmt == MFFT_REAL_F64, op->H, d.filters * d.channels, d.fh, d.fw, op->Ph,
                                op->Pw, d.mode == MFFT_CONV_VALID_CORR, s) != hipSuccess)
        return cfail(MFFT_ERR_HIP, "pack filters");
    int rc = mfft_exec(op->fwd_filt, op->H, op->H, s);
    if (rc) return cfail(rc, "%s", mfft_last_error());
    op->have_filters = true;
    return MFFT_OK;
}

static int conv_run(mfft_conv2d op, const void *in, int in_fmt, void *out, int out_fmt, int fsel, void *stream)
{
    if (!op || !in || !out) return cfail(MFFT_ERR_INVALID, "null argument");
    if (!op->have_filters) return cfail(MFFT_ERR_INVALID, "filters not set");
    if (fsel >= op->d.filters) return cfail(MFFT_ERR_INVALID, "filter %d of %d", fsel, op->d.filters);
    if (op->mid16 && fsel < 0 && op->d.filters > 1)
        return cfail(MFFT_ERR_UNSUPPORTED, "fp16 pass boundaries: one filter per call (mfft_conv2d_forward_select)");
    hipStream_t s = (hipStream_t)stream;
    OpCall call(op, s);
    if (call.err != hipSuccess) return cfail(MFFT_ERR_HIP, "stream order: %s", hipGetErrorString(call.err));
    const mfft_conv2d_desc &d = op->d;
    int rc;
    if (mfft::plan_first_pass_fusable(op->fwd_img)) {
        // the forward transform's first pass reads the real images, zero-padded
        mfft::ExecFuse f;
        f.real_src = in;
        f.real_f64 = in_fmt == MFFT_REAL_F64;
        f.real_b = d.batch;
        f.real_c = d.channels;
        f.real_h = d.h;
        f.real_w = d.w;
        rc = mfft::exec_fused(op->fwd_img, op->X, op->X, op->pairs * d.channels, s, f);
    } else {
        // Y holds at least as many spectra as X (one channel whenever mid16)
        void *packed = op->mid16 ? (void *)op->Y : (void *)op->X;
        mfft::trace_launch(s, "conv_pack", (long long)d.batch * d.channels * d.h * d.w * (in_fmt == MFFT_REAL_F64 ? 8 : 4),
                           op->pairs * d.channels * (long long)op->Ph * op->Pw * spec_eb(op));
        if (mfft::conv_pack_pairs(in, in_fmt == MFFT_REAL_F64, packed, op->spec_fmt, d.batch, d.channels, d.h, d.w,
                                  op->Ph, op->Pw, s) != hipSuccess)
            return cfail(MFFT_ERR_HIP, "pack");
        rc = mfft_exec(op->fwd_img, packed, op->X, s);
    }
    if (rc) return cfail(rc, "%s", mfft_last_error());
    const long long P = (long long)op->Ph * op->Pw;
    const int nf = fsel >= 0 ? 1 : d.filters;
    const float2 *H = fsel >= 0 ? op->H + (long long)fsel * d.channels * P : op->H;
    float2 *res = op->Y;  // the inverse's output
    if (op->refq) {
        // the reference's method: channel sum in fp32, each output map quantised to fp16 with
        // its power-of-ten scale, unscaled fp16 inverse, dequantised by the crop
        mfft::trace_launch(s, "conv_pointwise", op->pairs * d.channels * P * 4 + (long long)nf * d.channels * P * 8,
                           op->pairs * nf * P * 8);
        if (mfft::conv_pointwise(op->X, H, op->Y, MFFT_FMT_C16, op->pairs, d.channels, nf, P, s, MFFT_FMT_C32) !=
            hipSuccess)
            return cfail(MFFT_ERR_HIP, "pointwise");
        mfft::trace_launch(s, "conv_quantise", 2 * op->pairs * nf * P * 8, op->pairs * nf * P * 4);
        if (mfft::conv_quantise(op->Y, op->Q, op->qpart, op->qscale, op->pairs * nf, P, op->qchunks, s) != hipSuccess)
            return cfail(MFFT_ERR_HIP, "quantise");
        rc = mfft_exec_ex(op->inv, op->Q, op->Q, op->pairs * nf, nullptr, s);
        if (rc) return cfail(rc, "%s", mfft_last_error());
        mfft::trace_launch(s, "conv_crop", op->pairs * nf * (long long)op->oh * op->ow * 4,
                           (long long)d.batch * nf * op->oh * op->ow * (out_fmt == MFFT_REAL_F64 ? 8 : 4));
        if (mfft::conv_crop(op->Q, MFFT_FMT_C16, out, out_fmt == MFFT_REAL_F64, d.batch, nf, op->oh, op->ow, op->Ph,
                            op->Pw, op->oy, op->ox, (float)(1.0 / (double)P), s, op->qscale) != hipSuccess)
            return cfail(MFFT_ERR_HIP, "crop");
        return MFFT_OK;
    }
    if (d.channels == 1 && nf == 1 && mfft::plan_first_pass_fusable(op->inv)) {
        // one filter, one channel: the inverse's first pass multiplies by the spectrum
        mfft::ExecFuse f;
        f.in_mul = H;
        f.in_mul_period = P;
        rc = mfft::exec_fused(op->inv, op->X, op->Y, op->pairs, s, f);
    } else {
        mfft::trace_launch(s, "conv_pointwise", op->pairs * d.channels * P * spec_eb(op) + (long long)nf * d.channels * P * 8,
                           op->pairs * nf * P * spec_eb(op));
        if (mfft::conv_pointwise(op->X, H, op->Y, op->spec_fmt, op->pairs, d.channels, nf, P, s) != hipSuccess)
            return cfail(MFFT_ERR_HIP, "pointwise");
        if (op->mid16) res = op->X;  // out of place; X holds pairs x 1 spectra and nf == 1 here
        rc = mfft_exec_ex(op->inv, op->Y, res, op->pairs * nf, nullptr, s);
    }
    if (rc) return cfail(rc, "%s", mfft_last_error());
    mfft::trace_launch(s, "conv_crop", op->pairs * nf * (long long)op->oh * op->ow * (op->res_fmt == MFFT_FMT_C16 ? 4 : 8),
                       (long long)d.batch * nf * op->oh * op->ow * (out_fmt == MFFT_REAL_F64 ? 8 : 4));
    if (mfft::conv_crop(res, op->res_fmt, out, out_fmt == MFFT_REAL_F64, d.batch, nf, op->oh, op->ow, op->Ph,
                        op->Pw, op->oy, op->ox, 1.0f, s) != hipSuccess)
        return cfail(MFFT_ERR_HIP, "crop");
    return MFFT_OK;
}

int mfft_conv2d_forward(mfft_conv2d op, const void *in, int in_fmt, void *out, int out_fmt, void *stream)
{
    return conv_run(op, in, in_fmt, out, out_fmt, -1, stream);
}

int mfft_conv2d_forward_select(mfft_conv2d op, const void *in, int in_fmt, void *out, int out_fmt, int filter_index,
                               void *stream)
{
    if (filter_index < 0) return cfail(MFFT_ERR_INVALID, "filter index");
    return conv_run(op, in, in_fmt, out, out_fmt, filter_index, stream);
}

int mfft_conv2d_forward_two(mfft_conv2d op, const void *in0, const void *in1, int in_fmt, void *out0, void *out1,
                            int out_fmt, int filter0, int filter1, void *stream)
{
    if (!op || !in0 || !in1 || !out0 || !out1) return cfail(MFFT_ERR_INVALID, "null argument");
    if (!op->have_filters) return cfail(MFFT_ERR_INVALID, "filters not set");
    const mfft_conv2d_desc &d = op->d;
    if (d.channels != 1) return cfail(MFFT_ERR_UNSUPPORTED, "forward_two needs a single-channel operator");
    if (op->refq) return cfail(MFFT_ERR_UNSUPPORTED, "forward_two: not in the reference-method mode");
    if (filter0 < 0 || filter1 < 0 || filter0 >= d.filters || filter1 >= d.filters)
        return cfail(MFFT_ERR_INVALID, "filter index");
    hipStream_t s = (hipStream_t)stream;
    OpCall call(op, s);
    if (call.err != hipSuccess) return cfail(MFFT_ERR_HIP, "stream order: %s", hipGetErrorString(call.err));
    const long long P = (long long)op->Ph * op->Pw;
    int rc;
    if (mfft::plan_first_pass_fusable(op->fwd_img)) {
        mfft::ExecFuse f;  // the two images read straight into the first pass
        f.real_src = in0;
        f.real_src2 = in1;
        f.real_f64 = in_fmt == MFFT_REAL_F64;
        f.real_b = 2;
        f.real_c = 1;
        f.real_h = d.h;
        f.real_w = d.w;
        rc = mfft::exec_fused(op->fwd_img, op->X, op->X, 1, s, f);
    } else {
        void *packed = op->mid16 ? (void *)op->Y : (void *)op->X;
        mfft::trace_launch(s, "conv_pack_two", 2LL * d.h * d.w * (in_fmt == MFFT_REAL_F64 ? 8 : 4), P * spec_eb(op));
        if (mfft::conv_pack_two(in0, in1, in_fmt == MFFT_REAL_F64, packed, op->spec_fmt, d.h, d.w, op->Ph, op->Pw, s) !=
            hipSuccess)
            return cfail(MFFT_ERR_HIP, "pack");
        rc = mfft_exec_ex(op->fwd_img, packed, op->X, 1, nullptr, s);
    }
    if (rc) return cfail(rc, "%s", mfft_last_error());
    mfft::trace_launch(s, "conv_pointwise_two", P * spec_eb(op) + 2 * P * 8, P * spec_eb(op));
    if (mfft::conv_pointwise_two(op->X, op->H + filter0 * P, op->H + filter1 * P, op->Y, op->spec_fmt, op->Ph, op->Pw,
                                 s) != hipSuccess)
        return cfail(MFFT_ERR_HIP, "pointwise");
    float2 *res = op->mid16 ? op->X : op->Y;  // mid16: the inverse runs out of place
    rc = mfft_exec_ex(op->inv, op->Y, res, 1, nullptr, s);
    if (rc) return cfail(rc, "%s", mfft_last_error());
    mfft::trace_launch(s, "conv_crop_two", (long long)op->oh * op->ow * (op->res_fmt == MFFT_FMT_C16 ? 4 : 8),
                       2LL * op->oh * op->ow * (out_fmt == MFFT_REAL_F64 ? 8 : 4));
    if (mfft::conv_crop_two(res, op->res_fmt, out0, out1, out_fmt == MFFT_REAL_F64, op->oh, op->ow, op->Pw, op->oy,
                            op->ox, s) != hipSuccess)
        return cfail(MFFT_ERR_HIP, "crop");
    return MFFT_OK;
}

// Gradients of the 'valid' correlation layer (the reference's conv2d_backward,
// tensorcore_fft_conv2d.cpp:364-442): with delta = dL/dout [B][F][oh][ow],
//   grad_w[f][c]  = sum_b in[b][c] (correlated with) delta[b][f]   -> IFFT(X conj(D)), window [0,fh)x[0,fw)
//   grad_in[b][c] = sum_f delta[b][f] (full conv) w[f][c]          -> IFFT(D conj(H_rot)), read at
//                   the cyclic offset -(fh-1), -(fw-1) (H_rot: spectrum of the rotated filter)
// on the forward's padded size (Ph >= h: neither product wraps into the window read).
static int backward_setup(mfft_conv2d op)
{
    if (op->fwd_delta) return MFFT_OK;
    const mfft_conv2d_desc &d = op->d;
    mfft_desc fd;
    mfft_desc_default(&fd);
    fd.precision = d.precision;
    fd.in_fmt = fd.out_fmt = MFFT_FMT_C32;
    fd.workspace_limit = d.workspace_limit;
    fd.scale = MFFT_SCALE_STAGE;
    int rc = mfft_plan_create_2d(&op->fwd_delta, op->Ph, op->Pw, op->pairs * d.filters, &fd);
    fd.scale = MFFT_SCALE_NONE;
    fd.direction = MFFT_INVERSE;
    if (!rc) rc = mfft_plan_create_2d(&op->inv_fc, op->Ph, op->Pw, (int64_t)d.filters * d.channels, &fd);
    if (!rc) rc = mfft_plan_create_2d(&op->inv_pc, op->Ph, op->Pw, op->pairs * d.channels, &fd);
    if (!rc && hipMalloc(&op->G, (size_t)d.filters * d.channels * op->Ph * op->Pw * 8) != hipSuccess) rc = MFFT_ERR_NOMEM;
    if (rc) {
        std::string m = mfft_last_error();
        for (mfft_plan *p : {&op->fwd_delta, &op->inv_fc, &op->inv_pc})
            if (*p) {
                mfft_plan_destroy(*p);
                *p = nullptr;
            }
        return cfail(rc, "backward setup: %s", m.c_str());
    }
    return MFFT_OK;
}

int mfft_conv2d_backward(mfft_conv2d op, const void *in, const void *delta, int real_fmt, void *grad_in, void *grad_w,
                         void *stream)
{
    if (!op || !delta || (!grad_in && !grad_w) || (grad_w && !in)) return cfail(MFFT_ERR_INVALID, "null argument");
    if (op->d.mode != MFFT_CONV_VALID_CORR) return cfail(MFFT_ERR_UNSUPPORTED, "backward of the 'valid' correlation mode only");
    if (op->d.spec_fp16 != MFFT_CONV_SPEC_FP32)
        return cfail(MFFT_ERR_UNSUPPORTED, "backward needs fp32 spectra and intermediates (spec_fp16 = 0)");
    if (grad_in && !op->have_filters) return cfail(MFFT_ERR_INVALID, "filters not set");
    hipStream_t s = (hipStream_t)stream;
    OpCall call(op, s);
    if (call.err != hipSuccess) return cfail(MFFT_ERR_HIP, "stream order: %s", hipGetErrorString(call.err));
    int rc = backward_setup(op);
    if (rc) return rc;
    const mfft_conv2d_desc &d = op->d;
    const int f64 = real_fmt == MFFT_REAL_F64;
    const long long P = (long long)op->Ph * op->Pw;
    // delta spectra (scaled 1/P) in Y
    mfft::trace_launch(s, "conv_pack", (long long)d.batch * d.filters * op->oh * op->ow * (f64 ? 8 : 4),
                       op->pairs * d.filters * P * 8);
    if (mfft::conv_pack_pairs(delta, f64, op->Y, MFFT_FMT_C32, d.batch, d.filters, op->oh, op->ow, op->Ph, op->Pw, s) !=
        hipSuccess)
        return cfail(MFFT_ERR_HIP, "pack delta");
    rc = mfft_exec(op->fwd_delta, op->Y, op->Y, s);
    if (rc) return cfail(rc, "%s", mfft_last_error());
    if (grad_w) {
        mfft::trace_launch(s, "conv_pack", (long long)d.batch * d.channels * d.h * d.w * (f64 ? 8 : 4),
                           op->pairs * d.channels * P * 8);
        if (mfft::conv_pack_pairs(in, f64, op->X, MFFT_FMT_C32, d.batch, d.channels, d.h, d.w, op->Ph, op->Pw, s) !=
            hipSuccess)
            return cfail(MFFT_ERR_HIP, "pack input");
        rc = mfft_exec(op->fwd_img, op->X, op->X, s);
        if (rc) return cfail(rc, "%s", mfft_last_error());
        mfft::trace_launch(s, "conv_pointwise_gw", op->pairs * (d.channels + d.filters) * P * 8,
                           (long long)d.filters * d.channels * P * 8);
        if (mfft::conv_pointwise_gw(op->X, op->Y, op->G, (int)op->pairs, d.channels, d.filters, op->Ph, op->Pw, s) !=
            hipSuccess)
            return cfail(MFFT_ERR_HIP, "pointwise (filter gradient)");
        rc = mfft_exec(op->inv_fc, op->G, op->G, s);
        if (rc) return cfail(rc, "%s", mfft_last_error());
        // both spectra carry 1/P and the unnormalised inverse one factor P back
        mfft::trace_launch(s, "conv_crop_real", (long long)d.filters * d.channels * d.fh * d.fw * 8,
                           (long long)d.filters * d.channels * d.fh * d.fw * (f64 ? 8 : 4));
        if (mfft::conv_crop_real(op->G, grad_w, f64, (long long)d.filters * d.channels, d.fh, d.fw, op->Ph, op->Pw,
                                 (float)P, s) != hipSuccess)
            return cfail(MFFT_ERR_HIP, "crop (filter gradient)");
    }
    if (grad_in) {
        mfft::trace_launch(s, "conv_pointwise_gi", op->pairs * d.filters * P * 8 + (long long)d.filters * d.channels * P * 8,
                           op->pairs * d.channels * P * 8);
        if (mfft::conv_pointwise_gi(op->Y, op->H, op->X, op->pairs, d.channels, d.filters, P, s) != hipSuccess)
            return cfail(MFFT_ERR_HIP, "pointwise (input gradient)");
        rc = mfft_exec(op->inv_pc, op->X, op->X, s);
        if (rc) return cfail(rc, "%s", mfft_last_error());
        mfft::trace_launch(s, "conv_crop_cyclic", op->pairs * d.channels * (long long)d.h * d.w * 8,
                           (long long)d.batch * d.channels * d.h * d.w * (f64 ? 8 : 4));
        if (mfft::conv_crop_cyclic(op->X, grad_in, f64, d.batch, d.channels, d.h, d.w, op->Ph, op->Pw,
                                   op->Ph - (d.fh - 1), op->Pw - (d.fw - 1), 1.0f, s) != hipSuccess)
            return cfail(MFFT_ERR_HIP, "crop (input gradient)");
    }
    return MFFT_OK;
}

int mfft_conv2d_info_get(mfft_conv2d op, mfft_conv2d_info *info)
{
    if (!op || !info) return cfail(MFFT_ERR_INVALID, "null argument");
    info->pad_h = op->Ph;
    info->pad_w = op->Pw;
    info->out_h = op->oh;
    info->out_w = op->ow;
    info->device_bytes = op->bytes;
    info->spectrum_bytes = op->spec_bytes;
    return MFFT_OK;
}

}  // extern "C"
