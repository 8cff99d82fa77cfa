// Fast tile kernels, fp16 and bf16 precision (instantiations).
#include "mfft_fast_impl.h"
namespace mfft {
fast_fn find_fast_half(int prec, int R, int kind, int inf, int outf)
{
    using namespace fast;
    if (prec == PREC_FP16 && inf == FMT_C16 && outf == FMT_C16) return pick_kind<PREC_FP16, FMT_C16, FMT_C16>(R, kind);
    if (prec == PREC_BF16 && inf == FMT_CBF16 && outf == FMT_CBF16)
        return pick_kind<PREC_BF16, FMT_CBF16, FMT_CBF16>(R, kind);
    return nullptr;
}
}  // namespace mfft
