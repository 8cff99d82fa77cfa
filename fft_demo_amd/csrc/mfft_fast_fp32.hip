// Fast tile kernels, plain fp32 vector-ALU precision (the comparator; instantiations).
#include "mfft_fast_impl.h"
namespace mfft {
fast_fn find_fast_split(int R, int kind, int inf, int outf);
fast_fn find_fast_half(int prec, int R, int kind, int inf, int outf);

fast_fn find_fast_kernel(int prec, int R, int kind, int inf, int outf)
{
    using namespace fast;
    if (prec == PREC_SPLIT) return find_fast_split(R, kind, inf, outf);
    if (prec == PREC_FP16 || prec == PREC_BF16) return find_fast_half(prec, R, kind, inf, outf);
    if (prec == PREC_FP32 && inf == FMT_C32 && outf == FMT_C32) return pick_kind<PREC_FP32, FMT_C32, FMT_C32>(R, kind);
    return nullptr;
}

int fast_tile_lines(int R, int kind) { return kind == KIND_COLS ? cols_lines(R) : kMaxTile / R; }
int fast_threads(int prec, int R, int kind) { return fast_waves_of(prec, R, kind) * 64; }
}  // namespace mfft
