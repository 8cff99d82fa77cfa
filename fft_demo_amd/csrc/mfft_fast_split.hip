// Fast tile kernels, split-fp16 precision (instantiations). The file is compiled once per
// (input, output) format pair, MFFT_PART = 0..3, so the four objects build in parallel;
// part 0 also holds the dispatcher.
#include "mfft_fast_impl.h"

#ifndef MFFT_PART
#error "compile with -DMFFT_PART=0..3"
#endif

namespace mfft {
#if MFFT_PART == 0
static fast_fn pick_part(int R, int kind) { return fast::pick_kind<PREC_SPLIT, FMT_C16, FMT_C32>(R, kind); }
#elif MFFT_PART == 1
static fast_fn pick_part(int R, int kind) { return fast::pick_kind<PREC_SPLIT, FMT_C32, FMT_C16>(R, kind); }
#elif MFFT_PART == 2
static fast_fn pick_part(int R, int kind) { return fast::pick_kind<PREC_SPLIT, FMT_C16, FMT_C16>(R, kind); }
#else
static fast_fn pick_part(int R, int kind) { return fast::pick_kind<PREC_SPLIT, FMT_C32, FMT_C32>(R, kind); }
#endif

#define MFFT_PART_FN_(k) find_fast_split_part##k
#define MFFT_PART_FN(k) MFFT_PART_FN_(k)
fast_fn MFFT_PART_FN(MFFT_PART)(int R, int kind) { return pick_part(R, kind); }

#if MFFT_PART == 0
fast_fn find_fast_split_part1(int R, int kind);
fast_fn find_fast_split_part2(int R, int kind);
fast_fn find_fast_split_part3(int R, int kind);

fast_fn find_fast_split(int R, int kind, int inf, int outf)
{
    if (inf == FMT_C16 && outf == FMT_C32) return find_fast_split_part0(R, kind);
    if (inf == FMT_C32 && outf == FMT_C16) return find_fast_split_part1(R, kind);
    if (inf == FMT_C16 && outf == FMT_C16) return find_fast_split_part2(R, kind);
    if (inf == FMT_C32 && outf == FMT_C32) return find_fast_split_part3(R, kind);
    return nullptr;
}
#endif
}  // namespace mfft
