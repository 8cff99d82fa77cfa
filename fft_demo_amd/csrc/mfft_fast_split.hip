// Fast tile kernels, split-fp16 precision (instantiations).
#include "mfft_fast_impl.h"
namespace mfft {
fast_fn find_fast_split(int R, int kind, int inf, int outf)
{
    using namespace fast;
    if (inf == FMT_C16 && outf == FMT_C32) return pick_kind<PREC_SPLIT, FMT_C16, FMT_C32>(R, kind);
    if (inf == FMT_C32 && outf == FMT_C16) return pick_kind<PREC_SPLIT, FMT_C32, FMT_C16>(R, kind);
    if (inf == FMT_C16 && outf == FMT_C16) return pick_kind<PREC_SPLIT, FMT_C16, FMT_C16>(R, kind);
    if (inf == FMT_C32 && outf == FMT_C32) return pick_kind<PREC_SPLIT, FMT_C32, FMT_C32>(R, kind);
    return nullptr;
}
}  // namespace mfft
