// Fast tile kernels, fp16 and bf16 precision (instantiations). Compiled once per precision,
// MFFT_PART = 0 (fp16, with the dispatcher) and 1 (bf16), so the two objects build in
// parallel.
#include "mfft_fast_impl.h"

#ifndef MFFT_PART
#error "compile with -DMFFT_PART=0..1"
#endif

namespace mfft {
#if MFFT_PART == 0
fast_fn find_fast_bf16(int R, int kind);

fast_fn find_fast_half(int prec, int R, int kind, int inf, int outf)
{
    using namespace fast;
    if (prec == PREC_FP16 && inf == FMT_C16 && outf == FMT_C16) return pick_kind<PREC_FP16, FMT_C16, FMT_C16>(R, kind);
    if (prec == PREC_BF16 && inf == FMT_CBF16 && outf == FMT_CBF16) return find_fast_bf16(R, kind);
    return nullptr;
}
#else
fast_fn find_fast_bf16(int R, int kind) { return fast::pick_kind<PREC_BF16, FMT_CBF16, FMT_CBF16>(R, kind); }
#endif
}  // namespace mfft
