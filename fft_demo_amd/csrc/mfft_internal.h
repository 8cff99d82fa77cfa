// Library-internal entry points shared between the planner and the conv operator (not part
// of the C ABI).
#pragma once
#include "../../include/mfft.h"
#include "mfft_common.h"

#include <vector>

namespace mfft {

// Work fused into the first pass of a plan (the streaming short-DFT kernel only): an
// elementwise multiplier on the input (the conv pointwise product) and/or an input built
// from real images, zero-padded (the conv packing). See TileArgs.
struct ExecFuse {
    const float2 *in_mul = nullptr;
    long long in_mul_period = 0;
    const void *real_src = nullptr;
    const void *real_src2 = nullptr;  // imaginary part from a second single image
    int real_f64 = 0, real_b = 0, real_c = 0, real_h = 0, real_w = 0;
};

// mfft_plan_create_2d whose passes keep the intermediates they store in the output buffer in
// the desc's mid_fmt (only the last pass writes out_fmt there): 16-bit pass boundaries for a
// plan with 32-bit input and output. Its execs need out != in unless the first pass reads a
// fused real source instead of the input.
int plan_create_2d_out_mid(mfft_plan *plan, int nx, int ny, int64_t batch, const mfft_desc *d);

// Whether the plan's first pass can take an ExecFuse.
bool plan_first_pass_fusable(mfft_plan p);
// mfft_exec_ex with the fused first pass (no caller workspace); MFFT_ERR_UNSUPPORTED when
// the first pass is not fusable.
int exec_fused(mfft_plan p, const void *in, void *out, int64_t batch, void *stream, const ExecFuse &f);

// Set the calling thread's message returned by mfft_last_error().
void set_last_error(const char *msg);

// Launch trace (mfft_trace_begin/end): every launch site calls this just before its
// kernel launch on stream s, with the kernel family and its algorithmic bytes. No-op
// unless the calling thread has a trace active.
void trace_launch(hipStream_t s, const char *kernel, long long bytes_read, long long bytes_written);
bool trace_active();
const char *fmt_name(int fmt);  // "c16", "cbf16", "c32"

// The device address, on the current device, of the array mfft_exec_noalloc uses as its
// workspace (resolved with hipGetSymbolAddress on first use per device, then cached).
int noalloc_scratch(void **out);

// Orders the streams that use one device buffer (a plan's workspace, a conv operator's
// spectra). Calls are made under the owner's mutex: acquire(s) before enqueuing work on s
// that touches the buffer, release(s) after. While every call uses one stream nothing is
// recorded (no cost on the common path). The first time another stream shows up the
// device is synchronised once (the previous stream may no longer exist, so no event can
// be recorded on it after the fact); from then on every release records an event that the
// next acquire on a different stream waits for. Work being captured into a graph is not
// tracked: graph launches are ordered by the caller's stream.
struct StreamOrder {
    hipStream_t last = nullptr;
    bool used = false;
    hipEvent_t ev = nullptr;  // non-null once several streams have been seen
    hipError_t acquire(hipStream_t s);
    hipError_t release(hipStream_t s);
    void destroy();
};

// The radix list whose mixed-radix digit reversal (gen_rev, tcfft_half.cu:1351-1374) is the
// input order of MFFT_ORDER_DIGIT_REVERSED plans of length n: the reference's lists for the
// sizes tcfftCreate handles (tcfft_half.cu:1250-1330), else radix-16 digits, the 2/4/8 tail,
// then the odd factors.
void digit_reversal_radices(int n, std::vector<int> &out);

}  // namespace mfft
