// Library-internal entry points shared between the planner and the conv operator (not part
// of the C ABI).
#pragma once
#include "../../include/mfft.h"
#include "mfft_common.h"

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

// Whether the plan's first pass can take an ExecFuse.
bool plan_first_pass_fusable(mfft_plan p);
// mfft_exec_ex with the fused first pass (no caller workspace); MFFT_ERR_UNSUPPORTED when
// the first pass is not fusable.
int exec_fused(mfft_plan p, const void *in, void *out, int64_t batch, void *stream, const ExecFuse &f);

// Set the calling thread's message returned by mfft_last_error().
void set_last_error(const char *msg);

}  // namespace mfft
