// mfft_exec_noalloc: the plan exec with a library-held workspace, for callers that cannot
// allocate device memory (code under hipGraph capture, tight serving loops, harnesses that
// forbid hipMalloc). It has no counterpart in the reference.
//
// The workspace is `scratch`, a file-scope device array of MFFT_NOALLOC_SCRATCH_BYTES: each
// GPU the library runs on holds its own copy from the time the library's code object is
// loaded there until the process ends. In host code the name `scratch` is a host-side
// handle, not a device address; its address on a device is resolved once per device with
// hipGetSymbolAddress (with that device current) and only that pointer is passed on.
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU (different
// streams, no ordering between them) race on it. Nothing in it has to start at zero: each
// pass of an exec writes the intermediates the next pass reads, and the array keeps
// whatever the last call left.
#include "../../include/mfft.h"
#include "mfft_internal.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

static __device__ __align__(256) unsigned char scratch[MFFT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices];  // address of `scratch` on each device, null until resolved

int nfail(int code, const char *msg)
{
    mfft::set_last_error(msg);
    return code;
}

}  // namespace

namespace mfft {

// The device address of `scratch` on the current device, resolved on first use there.
int noalloc_scratch(void **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nfail(MFFT_ERR_HIP, "mfft_exec_noalloc: hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices) return nfail(MFFT_ERR_UNSUPPORTED, "mfft_exec_noalloc: device index over 63");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char buf[160];
            snprintf(buf, sizeof buf, "mfft_exec_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return nfail(MFFT_ERR_HIP, buf);
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return MFFT_OK;
}

}  // namespace mfft

extern "C" int mfft_exec_noalloc(mfft_plan plan, const void *in, void *out, int64_t batch, void *stream)
{
    if (!plan || !in || !out) return nfail(MFFT_ERR_INVALID, "mfft_exec_noalloc: null argument");
    mfft_info inf;
    int rc = mfft_plan_info(plan, &inf);
    if (rc) return rc;
    if (inf.workspace_bytes > MFFT_NOALLOC_SCRATCH_BYTES) {
        char buf[256];
        snprintf(buf, sizeof buf,
                 "mfft_exec_noalloc: the plan's intermediates take %lld bytes per batch chunk, more than the %lld "
                 "bytes of its scratch (create the plan with workspace_limit <= MFFT_NOALLOC_SCRATCH_BYTES)",
                 (long long)inf.workspace_bytes, (long long)MFFT_NOALLOC_SCRATCH_BYTES);
        return nfail(MFFT_ERR_NOMEM, buf);
    }
    void *ws = nullptr;
    rc = mfft::noalloc_scratch(&ws);
    if (rc) return rc;
    return mfft_exec_ex(plan, in, out, batch, ws, stream);
}
