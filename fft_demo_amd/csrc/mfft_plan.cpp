// Host side of the matrix-core FFT: planner, device tables, exec and the C ABI.
//
// The reference plans a fixed radix/"merging" table per size (tcfftCreate,
// tcfft_half.cu:1235-1345; tcfftCreate_2d, tcfft_half_2d.cu:1151-1221) and launches one
// in-block pass plus strided merge passes (tcfftExec, tcfft_half.cu:1177-1233). This
// planner decomposes any supported axis length with the four-step method instead:
//   L = L1 * L2 :  pass A = L2 strided DFT_L1 (stride L2) times the twiddle W_L^(n2*k1),
//                  then DFT_L2 on each row k1, written to k1 + L1*k2 (recursively).
// Every pass is one launch of the tile kernel; intermediates live in a per-plan
// workspace, processed in batch chunks small enough to stay in the 256 MB Infinity Cache.
#include "../../include/mfft.h"
#include "mfft_common.h"
#include "mfft_fast.h"
#include "mfft_internal.h"
#include "mfft_onchip.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace mfft {
hipError_t gen_twiddle(float2 *out, long long n, long long count, long long mult, int sign, hipStream_t s,
                       double scale = 1.0);
hipError_t gen_dft_ops(DftOps *ops, int sign, hipStream_t s);
hipError_t gen_w16(float *host3);
typedef void (*tile_fn)(TileArgs);
tile_fn find_tile_kernel(int prec, int R, int TL, bool generic);
int pow2_tile_lines(int R, int TL);
hipError_t launch_tile(const void *fn, const TileArgs &a, long long ntiles, hipStream_t stream, hipEvent_t start = nullptr,
                       hipEvent_t stop = nullptr);
const void *find_small_kernel(int R, int inf, int outf);
hipError_t launch_queue(int prec, const TileArgs *jobs, const long long *first_tile, int njobs, long long total,
                        unsigned long long *counter, int grid, hipStream_t s);
int queue_grid(int prec);
hipError_t stage_convert(const void *src, int sfmt, void *dst, int dfmt, long long n_inner, long long batch,
                         const int *perm_x, int nx, const int *perm_y, int ny, double scale, hipStream_t s);
}  // namespace mfft

using namespace mfft;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(MFFT_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------ factorisation
static bool is_pow2(long long n) { return n > 0 && (n & (n - 1)) == 0; }

// Stage radices of one in-tile line length, tail stage first (matches fft_tile_pow2).
static bool line_radices(int R, std::vector<int> &out)
{
    out.clear();
    if (R < 2) return false;
    int n = R, e2 = 0;
    while (n % 2 == 0) { n /= 2; ++e2; }
    int tail2 = 1 << (e2 % 4);
    int n16 = e2 / 4;
    if (tail2 > 1) out.push_back(tail2);
    // odd factors
    const int odd[] = {17, 5, 3};
    for (int p : odd)
        while (n % p == 0) { out.push_back(p); n /= p; }
    if (n != 1) return false;
    for (int i = 0; i < n16; ++i) out.push_back(16);
    return true;
}

// The radix list that defines the digit-reversed input order of a length-n axis.
// The sizes tcfftCreate knows get the reference's own lists (tcfft_half.cu:1250-1330:
// the tail after every second radix-16 digit, where its 512/1024-point passes put it);
// every other length: radix-16 digits first, then the 2/4/8 tail, then odd factors.
static void rev_radices(int n, std::vector<int> &out)
{
    out.clear();
    switch (n) {
    case 1 << 17: out = {16, 16, 2, 16, 16}; return;
    case 1 << 18: out = {16, 16, 2, 16, 16, 2}; return;
    case 1 << 19: out = {16, 16, 4, 16, 16, 2}; return;
    case 1 << 24: out = {16, 16, 16, 16, 16, 16}; return;
    case 1 << 25: out = {16, 16, 2, 16, 16, 16, 16}; return;
    case 1 << 26: out = {16, 16, 2, 16, 16, 2, 16, 16}; return;
    case 1 << 27: out = {16, 16, 2, 16, 16, 2, 16, 16, 2}; return;
    default: break;
    }
    int m = n, e2 = 0;
    while (m % 2 == 0) { m /= 2; ++e2; }
    for (int i = 0; i < e2 / 4; ++i) out.push_back(16);
    if (e2 % 4) out.push_back(1 << (e2 % 4));
    const int odd[] = {17, 5, 3};
    for (int p : odd)
        while (m % p == 0) { out.push_back(p); m /= p; }
}

static void gen_rev(int n, const std::vector<int> &radices, std::vector<int> &rev)
{
    // restates gen_rev (tcfft_half.cu:1351-1374)
    std::vector<int> a(n), b(n);
    for (int i = 0; i < n; ++i) a[i] = i;
    int now = n;
    for (int i = (int)radices.size() - 1; i >= 0; --i) {
        int r = radices[i], sub = now / r;
        for (int j = 0; j < n; j += now)
            for (int k = 0; k < r; ++k)
                for (int l = 0; l < sub; ++l) b[j + l + k * sub] = a[j + l * r + k];
        now = sub;
        std::swap(a, b);
    }
    rev = a;
}

// DigitRev of an index whose digits, least significant first, have the radices `digits`
// (powers of two, at most four digits): the result holds the digits in reverse order.
static bool make_digit_rev(const std::vector<int> &digits, DigitRev &d)
{
    memset(&d, 0, sizeof d);
    if (digits.size() > 4) return false;
    int total = 0;
    for (int r : digits) {
        if (!is_pow2(r)) return false;
        total += __builtin_ctz(r);
    }
    int sh = 0;
    d.n = (int)digits.size();
    for (size_t i = 0; i < digits.size(); ++i) {
        const int w = __builtin_ctz(digits[i]);
        d.sh[i] = sh;
        d.mask[i] = digits[i] - 1;
        d.out[i] = total - sh - w;
        sh += w;
    }
    return true;
}

// Lines of length R stored in the digit-reversed order of `list` (position p holds natural
// point gen_rev(list)[p]) can be read by the fast kernels' first stage when the list starts
// with 16 and the rest reverses with a DigitRev q: natural point col + (R/16) j is then at
// position j + 16 q(col). Fills q and checks that identity on every point.
static bool fast_rev_spec(int R, const std::vector<int> &list, DigitRev &q)
{
    if (R < 64 || !is_pow2(R) || list.empty() || list[0] != 16) return false;
    long long prod = 1;
    for (int r : list) prod *= r;
    if (prod != R) return false;
    std::vector<int> col_digits(list.rbegin(), list.rend() - 1);  // r_{m-1}, ..., r_1
    if (!make_digit_rev(col_digits, q)) return false;
    std::vector<int> rev;
    gen_rev(R, list, rev);
    const int R16 = R / 16;
    for (int pos = 0; pos < R; ++pos) {
        const int n = rev[pos];
        if (pos != n / R16 + 16 * digit_rev(n % R16, q)) return false;
    }
    return true;
}

// ------------------------------------------------------------------ plan structures
struct Level {
    long long count, in_stride, out_stride;
    bool batch;  // the batch level (chunked at exec time)
};

enum Buf { BUF_IN = 0, BUF_OUT = 1, BUF_WORK = 2 };

struct PassSpec {
    int R;
    long long ps_in, ps_out;
    std::vector<Level> lines;
    int tw_N = 0;
    int tw_level = -1;  // index into lines
    int src = BUF_IN, dst = BUF_OUT;
    int perm_axis = -1; // axis whose digit-reversal is fused into this pass's loads
    // the pass's input lines are stored in the digit-reversed order of this radix list
    // (gen_rev, tcfft_half.cu:1351-1374); empty = natural order
    std::vector<int> rev_list;
    DigitRev rev_tw{};  // four-step twiddle line factor of a digit-reversed first pass
    long long work_stride = 0;  // workspace elements per transform when padded (0 = dense)
    // derived
    std::vector<int> radix;
    bool generic = false;
    int fast_kind = -1;   // KIND_* of the fast kernel, -1 = tile kernel
    bool small = false;   // streaming short-DFT kernel (R <= 16, one thread per line)
    float2 *stage_tw_s = nullptr, *tw_lo_s = nullptr;
    float last_mul = 1.f, tail_mul = 1.f;
    long long max_grid = 0;  // persistent grid of the fast kernel
    int TL = 1, T = 1;
    int l0 = 0, l1 = -1, l2 = -1;  // indices into lines
    const void *fn = nullptr;
    float2 *tw_lo = nullptr, *tw_hi = nullptr;
    int tw_bits = 0;
    float2 *stage_tw = nullptr;
    int *perm = nullptr;
    bool rev_fast = false;  // the fast kernel reads the digit-reversed lines (rev_q)
    DigitRev rev_q{};
    bool onchip = false;    // the whole 65536-point line in one pass (fft_onchip)
    bool onchip_cols = false;  // 2048-point lines, 32 per workgroup (fft_onchip_cols)
    bool onchip_rows = false;  // ... contiguous lines stored transposed (its rows kind)
};

struct mfft_plan_s {
    int kind = 1;
    int n = 0, nx = 0, ny = 0;
    long long batch = 0;
    mfft_desc d{};
    long long total = 0;      // complex elements per transform
    int mid_fmt = FMT_C32;
    std::vector<PassSpec> passes;
    DftOps *ops = nullptr;
    void *work = nullptr;
    long long work_bytes = 0;
    long long chunk = 1;
    double out_scale = 1.0;
    std::vector<void *> allocs;
    std::vector<int> rev_axis[2];
    std::vector<int> rev_list_axis[2];  // the radix lists rev_axis was generated from
    std::mutex exec_mu;       // the workspace is per plan
    StreamOrder ws_order;     // streams that used the plan's own workspace
    // A plan whose passes include an on-chip kernel (one workgroup per line, or per 32
    // columns) keeps the plan of the same transform without it for execs of fewer than
    // small_below transforms, which would leave most of the device idle on the on-chip pass.
    mfft_plan_s *small = nullptr;
    long long small_below = 0;
    // Intermediates that a pass stores in the output buffer are kept in mid_fmt (the last
    // pass writes out_fmt there): with 16-bit intermediates and 32-bit output, every pass
    // boundary then moves 16-bit words. Execs must not alias a pass's source and destination
    // across a format change (checked at exec).
    bool out_mid = false;
};

// Element formats a pass reads and writes.
static int src_fmt(const mfft_plan_s *p, const PassSpec &ps)
{
    if (ps.src == BUF_IN) return p->d.in_fmt;
    if (ps.src == BUF_OUT) return p->out_mid ? p->mid_fmt : p->d.out_fmt;
    return p->mid_fmt;
}
static int dst_fmt(const mfft_plan_s *p, const PassSpec &ps)
{
    if (ps.dst == BUF_OUT) return p->out_mid && &ps != &p->passes.back() ? p->mid_fmt : p->d.out_fmt;
    if (ps.dst == BUF_IN) return p->d.in_fmt;
    return p->mid_fmt;
}

// Transforms (lines, or images of 2048-row 2D plans) an on-chip pass needs to occupy the
// device enough to beat the multi-pass plan: a workgroup per 65536-point line, 2/3 of the
// compute units (65536 x 256 on 256 CUs: 36 us against 53 us; at 128 lines the two-pass plan is
// faster); a workgroup per 32 columns of 2048 points, 3/4 of them (2048 x 2048: one image
// 49 us against 38 us, eight 129 us against 149 us).
static long long device_cus()
{
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || ncu < 1)
        ncu = 256;
    return ncu;
}
static long long onchip_min_lines() { return (2 * device_cus() + 2) / 3; }
static long long onchip_rows_min_lines(int l1)
{
    const long long blocks = std::max(1, l1 / kOnchipColsLines);
    return std::max(1LL, (3 * device_cus() / 4 + blocks - 1) / blocks);
}
static long long onchip_cols_min_images(int ny)
{
    const long long blocks = std::max(1, ny / kOnchipColsLines);
    return std::max(1LL, (3 * device_cus() / 4 + blocks - 1) / blocks);
}

static int fmt_bytes(int fmt) { return fmt == FMT_C32 ? 8 : 4; }

// ------------------------------------------------------------------ axis planning
// fast_cols: the pass may run on the strided matrix-core kernel (natural-order input), which
// holds 8192 / R lines of up to kColsMaxR points per tile; the tile kernels hold at most
// kMaxTile points, so a strided single pass there needs R <= kMaxTile / 16 for 16 lines.
static bool single_pass_ok(int R, long long ps_in, const std::vector<Level> &lines, bool fast_cols)
{
    if (R < 2 || R > kMaxTile) return false;
    std::vector<int> rad;
    if (!line_radices(R, rad)) return false;
    if (ps_in == 1) return true;
    bool has_unit = false;
    for (auto &l : lines)
        if (l.in_stride == 1) has_unit = true;
    if (!has_unit) return true;
    return R * 16 <= kMaxTile || (fast_cols && is_pow2(R) && R >= 64 && R <= kColsMaxR);
}

static bool factorable(long long n)
{
    std::vector<int> r;
    return n <= kMaxTile && line_radices((int)n, r);
}

// Append the passes of a length-L DFT along an axis. Input element (line, j) is at
// line_in + j*ps_in, output element (line, k) at line_out + k*ps_out.
static long long odd_part(long long L)
{
    while (L > 0 && L % 2 == 0) L /= 2;
    return L;
}

// Odd factors the streaming kernel takes as a pass of its own.
static bool small_odd(long long o) { return o == 3 || o == 5 || o == 9 || o == 15 || o == 17; }

// Whether a plan's intermediates between passes are 16-bit (the rule finalize_plan applies).
static bool narrow_mid(const mfft_desc *d)
{
    if (d->mid_fmt >= 0) return d->mid_fmt != MFFT_FMT_C32;
    return d->precision == MFFT_PREC_FP16 || d->precision == MFFT_PREC_BF16;
}

// narrow: the intermediates between passes are 16-bit (an extra pass costs 8 bytes a point)
// rev_cols: a digit-reversed strided axis may still run in one pass of the strided
// matrix-core kernel (it reads the reversed order itself)
static int plan_axis(std::vector<PassSpec> &out, long long L, long long ps_in, long long ps_out,
                     std::vector<Level> lines, int src, int dst, int perm_axis, int split_hint, bool narrow,
                     bool rev_cols = false)
{
    // odd * 2^a (2176 = 17 x 128): the odd factor in a streaming pass, the power of two on
    // the matrix-core kernels, instead of one generic mixed-radix tile pass.
    // split_hint = 1 keeps the single pass.
    const long long odd = odd_part(L);
    const bool odd_split = perm_axis < 0 && split_hint != 1 && odd > 1 && small_odd(odd) && L / odd >= 16 &&
                           L / odd <= kMaxTile;
    const bool fast_cols = (perm_axis < 0 || rev_cols) && split_hint <= 1;  // no split asked for
    if (!odd_split && single_pass_ok((int)std::min<long long>(L, 1 << 30), ps_in, lines, fast_cols)) {
        PassSpec p;
        p.R = (int)L;
        p.ps_in = ps_in;
        p.ps_out = ps_out;
        p.lines = lines;
        p.src = src;
        p.dst = dst;
        p.perm_axis = perm_axis;
        out.push_back(p);
        return MFFT_OK;
    }
    // pass A's destination: the workspace, or -- when this axis must end in the workspace
    // (rows of a 2D transform) -- the output buffer, which is free until the last axis
    int a_dst = BUF_WORK;
    if (dst == BUF_WORK) {
        if (!odd_split || src != BUF_IN)
            return fail(MFFT_ERR_UNSUPPORTED, "an inner axis longer than %d points is not supported", kMaxTile);
        a_dst = BUF_OUT;
    }
    if (perm_axis >= 0)
        return fail(MFFT_ERR_UNSUPPORTED,
                    "digit-reversed input is supported for axes done in one pass (length <= %d)", kMaxTile);
    // choose L1 (strided pass, <= 256) and L2 = L / L1
    long long L1 = 0;
    if (odd_split) {
        L1 = odd;
    } else if (split_hint > 1 && L % split_hint == 0 && factorable(split_hint) &&
               (split_hint <= 256 || (is_pow2(split_hint) && split_hint <= kColsMaxR))) {
        L1 = split_hint;
    } else if (is_pow2(L) && ps_in != 1 && L >= 512 && L <= 4096) {
        // strided axis (2D columns) of 512..4096 points: both passes are strided; the first
        // on the matrix-core kernel (L / 16 points, at least 64), the rest (16 or fewer) on
        // the streaming kernel (2048 x 2048: 1.1x the 16 + 128 split, measured)
        L1 = std::max<long long>(64, L / 16);
    } else if (is_pow2(L) && ps_in != 1 && L % 256 == 0 && L / 256 <= 256) {
        // longer strided axes: give the second pass the 256-point matrix-core kernel and the
        // small remainder to the first
        L1 = L / 256;
    } else if (is_pow2(L)) {
        int m = 0;
        while ((1LL << m) < L) ++m;
        int m1 = (m + 1) / 2;
        if (m1 > 8) m1 = 8;
        if (m - m1 > 12 && m1 < 8) m1 = std::min(8, m - 12);
        // 2^17..2^19 (contiguous lines): a long strided first pass -- 1024 points at 8 lines
        // per tile with 16-bit intermediates, 512 at 16 lines with 32-bit ones -- and the
        // rest (128..1024 points) in one pass; the autotuner's pick on MI355X for split fp16
        // and fp32, 1.02-1.18x the 256-point first pass (and the three passes below)
        if (ps_in == 1 && m >= 17 && m <= 19) {
            m1 = narrow ? 10 : 9;
        } else if (narrow && m >= 19 && m <= 21) {
            // 2^20..2^21 with 16-bit intermediates: a second pass of 2048 or 4096 points
            // would hold one or two lines per tile and store them transposed in 4- or
            // 8-byte pieces; three passes instead, the rest (8192 = 128 x 64) split again
            // below (split fp16: 1.05-1.3x faster; with fp32 intermediates the third pass
            // costs more than it saves)
            m1 = m - 13;
        }
        L1 = 1LL << m1;
    } else {
        for (long long c = 256; c >= 2; --c)
            if (L % c == 0 && factorable(c) && (L / c <= kMaxTile ? factorable(L / c) : true)) {
                L1 = c;
                break;
            }
    }
    if (L1 < 2 || L % L1 != 0) return fail(MFFT_ERR_UNSUPPORTED, "cannot factor axis length %lld", L);
    long long L2 = L / L1;
    // pass A: DFT_L1 over n1 (stride L2*ps_in) for each (line, n2); in-place layout
    PassSpec a;
    a.R = (int)L1;
    a.ps_in = L2 * ps_in;
    a.ps_out = L2 * ps_in;
    a.lines = lines;
    for (auto &l : a.lines) l.out_stride = l.in_stride;  // pass A keeps the input layout
    a.lines.push_back(Level{L2, ps_in, ps_in, false});
    a.tw_N = (int)L;
    a.tw_level = (int)a.lines.size() - 1;
    a.src = src;
    a.dst = a_dst;
    out.push_back(a);
    // remaining: DFT_L2 over n2 for each (line, k1); output at k1 + L1*k2
    // input side of the sub-problem reads pass A's output, laid out like the input
    std::vector<Level> sub2;
    for (auto &l : lines) sub2.push_back(Level{l.count, l.in_stride, l.out_stride, l.batch});
    sub2.push_back(Level{L1, L2 * ps_in, ps_out, false});
    return plan_axis(out, L2, ps_in, L1 * ps_out, sub2, a_dst, dst, -1, 1, narrow);
}

// 65536-point lines with 16-bit words in, out and between the two 256-point halves, in
// natural order or in the digit-reversed order of the radix list {16, 16, 16, 16}: one pass
// of the on-chip kernel (the line is read and written once) instead of the four-step pair,
// unless a split is asked for (split_hint > 1). The halves are the same four-step
// decomposition (tcfft_half.cu:266-390 then 69-161 for the reference's long lines).
static bool plan_onchip(mfft_plan_s *p, const mfft_desc *d)
{
    if (p->n != kOnchipN || d->split_hint > 1) return false;
    const bool rev = d->in_order == MFFT_ORDER_DIGIT_REVERSED;
    if (rev && p->rev_list_axis[0] != std::vector<int>{16, 16, 16, 16}) return false;
    const int fmt = d->precision == MFFT_PREC_BF16 ? MFFT_FMT_CBF16 : MFFT_FMT_C16;
    const int mid = d->mid_fmt >= 0                         ? d->mid_fmt
                    : d->precision == MFFT_PREC_FP16 ? MFFT_FMT_C16
                    : d->precision == MFFT_PREC_BF16 ? MFFT_FMT_CBF16
                                                     : MFFT_FMT_C32;
    if (d->in_fmt != fmt || d->out_fmt != fmt || mid != fmt) return false;
    if (!find_onchip_kernel(d->precision, fmt, rev)) return false;
    PassSpec ps;
    ps.R = kOnchipN;
    ps.ps_in = ps.ps_out = 1;
    ps.lines = {Level{p->batch, kOnchipN, kOnchipN, true}};
    ps.tw_N = kOnchipN;  // the four-step tables between the halves
    ps.tw_level = 0;
    ps.src = BUF_IN;
    ps.dst = BUF_OUT;
    ps.onchip = true;
    if (rev) ps.rev_list = p->rev_list_axis[0];
    p->passes.push_back(ps);
    return true;
}

// N = L1 x 2048 (2^19, 2^20) with 16-bit words in, out and between the passes: the four-step
// split with the 2048-point half on the on-chip line kernel -- pass A: DFT_L1 over n1 (stride
// 2048, the strided matrix-core kernel) times W_N^(n2 k1); pass B: DFT_2048 of each row k1, 32
// rows per workgroup read whole and stored transposed at k1 + L1 k2. Split fp16 on MI355X:
// 2^20 x 16 / 64 in 72 / 303 us against 78 / 316 us for the three-pass plan, 2^19 x 32 / 64 in
// 63 / 131 us against 68 / 172 us for 1024 + 512; at 2^17 and 2^18 (L1 = 64, 128) it was no
// faster than the 1024-point first pass, which those lengths keep.
static bool plan_onchip_rows(mfft_plan_s *p, const mfft_desc *d)
{
    const long long N = p->n;
    if (N < (1LL << 19) || N > (1LL << 20) || !is_pow2(N) || d->in_order != MFFT_ORDER_NATURAL || d->split_hint > 1)
        return false;
    const int L1 = (int)(N / kOnchipColsN);
    const int fmt = d->precision == MFFT_PREC_BF16 ? MFFT_FMT_CBF16 : MFFT_FMT_C16;
    const int mid = d->mid_fmt >= 0                         ? d->mid_fmt
                    : d->precision == MFFT_PREC_FP16 ? MFFT_FMT_C16
                    : d->precision == MFFT_PREC_BF16 ? MFFT_FMT_CBF16
                                                     : MFFT_FMT_C32;
    if (d->in_fmt != fmt || d->out_fmt != fmt || mid != fmt) return false;
    if (!find_onchip_cols_kernel(d->precision, fmt, true) || !find_fast_kernel(d->precision, L1, KIND_COLS, fmt, fmt))
        return false;
    PassSpec a;
    a.R = L1;
    a.ps_in = a.ps_out = kOnchipColsN;
    a.lines = {Level{p->batch, N, N, true}, Level{kOnchipColsN, 1, 1, false}};
    a.tw_N = (int)N;
    a.tw_level = 1;
    a.src = BUF_IN;
    a.dst = BUF_WORK;
    PassSpec b;
    b.R = kOnchipColsN;
    b.ps_in = 1;
    b.ps_out = L1;
    b.lines = {Level{L1, kOnchipColsN, 1, false}, Level{p->batch, N, N, true}};
    b.src = BUF_WORK;
    b.dst = BUF_OUT;
    b.onchip_cols = b.onchip_rows = true;
    p->passes.push_back(a);
    p->passes.push_back(b);
    return true;
}

// Digit-reversed input (tcfftExec's order, tcfft_half.cu:1177-1233) of a 1D length longer
// than one tile: N = L1 L2, split after the first m radices of the list. Position
// p = p_lo + L1 p_hi holds natural element n2 + L2 n1 with n1 = gen_rev(list[:m])[p_lo] and
// n2 = gen_rev(list[m:])[p_hi], so
//   pass A: each contiguous block p_hi is a digit-reversed line of list[:m]: DFT_L1 over n1,
//           times W_N^(n2 k1) (n2 = the digit reversal of p_hi), stored transposed at
//           k1 L2 + p_hi;
//   pass B: row k1 of the workspace is a digit-reversed line of list[m:] (position p_hi holds
//           n2): DFT_L2, stored transposed at k1 + L1 k2, natural order.
// Both passes read contiguous lines. The halves must start with a radix-16 digit and run
// on the fast kernels, and L1, L2 are kept to 256..1024 points (longer lines hold too few
// per tile for the transposed stores).
static int plan_reversed_1d(mfft_plan_s *p)
{
    const std::vector<int> &list = p->rev_list_axis[0];
    const long long N = p->n;
    int best = -1;
    double bestd = 1e9;
    long long L1 = 1;
    for (size_t m = 1; m < list.size(); ++m) {
        L1 *= list[m - 1];
        const long long L2 = N / L1;
        if (list[0] != 16 || list[m] != 16 || L1 < 256 || L1 > 1024 || L2 < 256 || L2 > 1024) continue;
        const double dd = std::fabs(std::log2((double)L1) - std::log2((double)L2));
        if (dd < bestd) {
            bestd = dd;
            best = (int)m;
        }
    }
    if (best < 0) return fail(MFFT_ERR_UNSUPPORTED, "digit-reversed input of length %lld in two passes", N);
    std::vector<int> lo(list.begin(), list.begin() + best), hi(list.begin() + best, list.end());
    L1 = 1;
    for (int r : lo) L1 *= r;
    const long long L2 = N / L1;
    PassSpec a;
    a.R = (int)L1;
    a.ps_in = 1;
    a.ps_out = L2;
    a.lines = {Level{L2, L1, 1, false}, Level{p->batch, N, N, true}};
    a.tw_N = (int)N;
    a.tw_level = 0;
    a.src = BUF_IN;
    a.dst = BUF_WORK;
    a.rev_list = lo;
    if (!make_digit_rev(hi, a.rev_tw)) return fail(MFFT_ERR_UNSUPPORTED, "twiddle line reversal");
    {
        std::vector<int> rh;
        gen_rev((int)L2, hi, rh);
        for (int ph = 0; ph < (int)L2; ++ph)
            if (rh[ph] != digit_rev(ph, a.rev_tw)) return fail(MFFT_ERR_UNSUPPORTED, "twiddle line reversal");
    }
    PassSpec b;
    b.R = (int)L2;
    b.ps_in = 1;
    b.ps_out = L1;
    b.lines = {Level{L1, L2, 1, false}, Level{p->batch, N, N, true}};
    b.src = BUF_WORK;
    b.dst = BUF_OUT;
    b.rev_list = hi;
    p->passes.push_back(a);
    p->passes.push_back(b);
    return MFFT_OK;
}

// Choose tile geometry and kernel for every pass.
static int finish_passes(mfft_plan_s *p)
{
    for (auto &ps : p->passes) {
        if (ps.onchip_cols) {
            ps.radix = {16, 16, 8};
            ps.fast_kind = KIND_ONCHIP_COLS;
            ps.fn = find_onchip_cols_kernel(p->d.precision, p->mid_fmt, ps.onchip_rows);
            ps.l0 = 0;
            ps.T = kOnchipColsLines;
            ps.TL = 1;
            if (!ps.fn) return fail(MFFT_ERR_UNSUPPORTED, "no on-chip column kernel for the formats");
            continue;
        }
        if (ps.onchip) {
            ps.radix = {16, 16, 16, 16};
            ps.fast_kind = KIND_ONCHIP;
            ps.fn = find_onchip_kernel(p->d.precision, p->d.in_fmt, !ps.rev_list.empty());
            ps.rev_fast = !ps.rev_list.empty();  // the kernel reads the stored order itself
            ps.l0 = 0;
            ps.T = 1;
            ps.TL = 1;
            if (!ps.fn) return fail(MFFT_ERR_UNSUPPORTED, "no on-chip kernel for the formats");
            continue;
        }
        if (!line_radices(ps.R, ps.radix)) return fail(MFFT_ERR_UNSUPPORTED, "line length %d", ps.R);
        if (ps.perm_axis >= 0 && ps.rev_list.empty()) ps.rev_list = p->rev_list_axis[ps.perm_axis];
        const bool rev = !ps.rev_list.empty();
        // merge adjacent levels that are contiguous on both sides (never the batch level)
        // choose l0
        int unit_in = -1, unit_out = -1;
        for (int i = 0; i < (int)ps.lines.size(); ++i) {
            if (ps.lines[i].in_stride == 1 && unit_in < 0) unit_in = i;
            if (ps.lines[i].out_stride == 1 && unit_out < 0) unit_out = i;
        }
        bool strided = ps.ps_in != 1;
        bool pow2 = is_pow2(ps.R) && ps.R >= 16;
        const int in_fmt = src_fmt(p, ps);
        const int out_fmt = dst_fmt(p, ps);
        if (p->out_mid && ps.src == ps.dst && in_fmt != out_fmt)
            return fail(MFFT_ERR_UNSUPPORTED, "in-place pass changes the element format");
        // digit-reversed lines on the fast kernels: the list starts with 16 and the rest
        // reverses digit by digit (every reference list does)
        ps.rev_fast = rev && fast_rev_ok(p->d.precision, in_fmt) && fast_rev_spec(ps.R, ps.rev_list, ps.rev_q);
        // strided lines longer than a 16-line tile kernel holds: the strided matrix-core
        // kernel when it exists for the formats (8192 / R lines of up to kColsMaxR points)
        const bool cols_long = strided && unit_in >= 0 && ps.R * 16 > kMaxTile && pow2 && ps.R <= kColsMaxR &&
                               (!rev || ps.rev_fast) && ps.lines[unit_in].out_stride == 1 &&
                               find_fast_kernel(p->d.precision, ps.R, KIND_COLS, in_fmt, out_fmt) != nullptr;
        if (strided && unit_in >= 0 && (ps.R * 16 <= kMaxTile || cols_long)) {
            ps.l0 = unit_in;
            ps.TL = 16;
        } else if (!strided) {
            ps.TL = 1;
            if (ps.ps_out != 1 && unit_out >= 0)
                ps.l0 = unit_out;
            else {
                // the level with the smallest input stride
                int best = 0;
                for (int i = 1; i < (int)ps.lines.size(); ++i)
                    if (ps.lines[i].in_stride < ps.lines[best].in_stride) best = i;
                ps.l0 = best;
            }
        } else {
            ps.l0 = 0;
            ps.TL = 1;
            pow2 = false;
        }
        ps.generic = !pow2;
        // short strided DFTs: streaming kernel, one thread per line
        if (!rev && ps.ps_in != 1 &&
            (ps.R == 2 || ps.R == 4 || ps.R == 8 || ps.R == 16 || small_odd(ps.R))) {
            const void *fn = find_small_kernel(ps.R, in_fmt, out_fmt);
            if (fn) {
                ps.small = true;
                ps.fn = fn;
                ps.l0 = unit_in >= 0 ? unit_in : 0;
                ps.T = 1;
                ps.TL = 1;
                ps.radix.assign(1, ps.R);
            }
        }
        // fast kernel: power of two 64..4096, natural order or digit-reversed lines it reads
        if (!ps.small && pow2 && ps.R >= 64 && ps.R <= kMaxTile && (!rev || ps.rev_fast)) {
            int kind = -1, l0 = -1;
            if (ps.ps_in == 1 && ps.ps_out == 1) {
                kind = KIND_ROWS;
                l0 = 0;
                for (int i = 1; i < (int)ps.lines.size(); ++i)
                    if (ps.lines[i].in_stride < ps.lines[l0].in_stride) l0 = i;
            } else if (ps.ps_in == 1 && unit_out >= 0) {
                kind = KIND_ROWS_TR;
                l0 = unit_out;
            } else if (ps.ps_in != 1 && ps.R <= kColsMaxR) {
                for (int i = 0; i < (int)ps.lines.size(); ++i)
                    if (ps.lines[i].in_stride == 1 && ps.lines[i].out_stride == 1) { l0 = i; break; }
                if (l0 >= 0) kind = KIND_COLS;
            }
            const void *fn = kind >= 0 ? (const void *)find_fast_kernel(p->d.precision, ps.R, kind, in_fmt, out_fmt) : nullptr;
            if (fn) {
                ps.fast_kind = kind;
                ps.l0 = l0;
                ps.T = fast_tile_lines(ps.R, kind);
                ps.TL = kind == KIND_COLS ? cols_lines(ps.R) : 1;
                ps.fn = fn;
                int dev = 0, ncu = 0, nb = 0;
                HIP_TRY(hipGetDevice(&dev));
                HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
                HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, fast_threads(p->d.precision, ps.R, kind), 0));
                ps.max_grid = (long long)std::max(1, nb) * std::max(1, ncu);
                ps.max_grid -= ps.max_grid % 8;
                if (ps.max_grid < 8) ps.max_grid = 8;
                // radix order of the fast schedule: 16, tail, 16, ...
                std::vector<int> r2;
                int tail = 1;
                for (int r : ps.radix) if (r != 16) tail = r;
                r2.push_back(16);
                if (tail > 1) r2.push_back(tail);
                for (size_t i = 1; i < ps.radix.size() - (tail > 1 ? 1 : 0); ++i) r2.push_back(16);
                ps.radix = r2;
            }
        }
        if (ps.fast_kind < 0) ps.rev_fast = false;
        if (rev && !ps.rev_fast && ps.perm_axis < 0)
            return fail(MFFT_ERR_UNSUPPORTED, "digit-reversed %d-point lines need the matrix-core kernels", ps.R);
        if (ps.fast_kind >= 0 || ps.small) {
            // l0 chosen above
        } else if (!ps.generic) {
            ps.T = pow2_tile_lines(ps.R, ps.TL);
        } else {
            int t = kMaxTile / ps.R;
            if (ps.TL == 16 && t < 16) ps.TL = 1;
            ps.T = std::max(1, t);
            if (ps.TL > ps.T) ps.TL = 1;
            if (ps.T % ps.TL) ps.T = ps.TL;
        }
        // remaining levels -> l1, l2 (at most two besides l0)
        std::vector<int> rest;
        if (ps.tw_N && ps.tw_level != ps.l0) rest.push_back(ps.tw_level);
        for (int i = 0; i < (int)ps.lines.size(); ++i)
            if (i != ps.l0 && !(ps.tw_N && i == ps.tw_level)) rest.push_back(i);
        if (rest.size() > 2) return fail(MFFT_ERR_UNSUPPORTED, "too many line levels");
        ps.l1 = rest.size() > 0 ? rest[0] : -1;
        ps.l2 = rest.size() > 1 ? rest[1] : -1;
        if (ps.tw_N && ps.tw_level != ps.l0 && ps.tw_level != ps.l1)
            return fail(MFFT_ERR_UNSUPPORTED, "twiddle level layout");
        if (ps.fast_kind < 0 && !ps.small) {
            ps.fn = (const void *)find_tile_kernel(p->d.precision, ps.R, ps.TL, ps.generic);
            if (!ps.fn) return fail(MFFT_ERR_UNSUPPORTED, "no kernel for R=%d TL=%d", ps.R, ps.TL);
        }
    }
    return MFFT_OK;
}

// The VALU butterflies' constants (TileArgs::w16), generated on the GPU once per process by
// the first plan or batcher.
static float g_w16[3];
static std::once_flag g_w16_once;
static hipError_t g_w16_err = hipSuccess;

static int ensure_w16()
{
    std::call_once(g_w16_once, [] { g_w16_err = gen_w16(g_w16); });
    if (g_w16_err != hipSuccess) return fail(MFFT_ERR_HIP, "generating the butterfly constants: %s", hipGetErrorString(g_w16_err));
    return MFFT_OK;
}

static int alloc_dev(mfft_plan_s *p, void **ptr, size_t bytes)
{
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(MFFT_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    p->allocs.push_back(*ptr);
    return MFFT_OK;
}

static int build_tables(mfft_plan_s *p)
{
    if (ensure_w16()) return MFFT_ERR_HIP;
    int sign = p->d.direction;
    if (alloc_dev(p, (void **)&p->ops, sizeof(DftOps))) return MFFT_ERR_NOMEM;
    HIP_TRY(gen_dft_ops(p->ops, sign, 0));
    for (auto &ps : p->passes) {
        // the on-chip kernel's stages are 256-point halves; the column kernel's second half
        // is 128 points (its 2048-point twiddle table is built below)
        const int Rst = ps.onchip ? 256 : ps.onchip_cols ? 128 : ps.R;
        if (alloc_dev(p, (void **)&ps.stage_tw, sizeof(float2) * Rst)) return MFFT_ERR_NOMEM;
        HIP_TRY(gen_twiddle(ps.stage_tw, Rst, Rst, 1, sign, 0));
        if (ps.tw_N) {
            int bits = 0;
            while ((1LL << (2 * bits)) < ps.tw_N) ++bits;
            long long nlo = 1LL << bits, nhi = (ps.tw_N + nlo - 1) / nlo;
            ps.tw_bits = bits;
            if (alloc_dev(p, (void **)&ps.tw_lo, sizeof(float2) * nlo)) return MFFT_ERR_NOMEM;
            if (alloc_dev(p, (void **)&ps.tw_hi, sizeof(float2) * nhi)) return MFFT_ERR_NOMEM;
            HIP_TRY(gen_twiddle(ps.tw_lo, ps.tw_N, nlo, 1, sign, 0));
            HIP_TRY(gen_twiddle(ps.tw_hi, ps.tw_N, nhi, nlo, sign, 0));
        }
        if (ps.fast_kind >= 0) {
            // radix-16 stage scale (1/16 per stage when scaling), divided by the factor 16
            // carried by the matrix-core operands; the fp32 path has no such factor
            double fac = (p->d.scale == MFFT_SCALE_STAGE ? 1.0 / 16 : 1.0) /
                         (p->d.precision == MFFT_PREC_FP32 ? 1.0 : 16.0);
            ps.last_mul = (float)fac;
            int tail = 1;
            for (int r : ps.radix) if (r != 16) tail = r;
            ps.tail_mul = (p->d.scale == MFFT_SCALE_STAGE) ? 1.0f / tail : 1.0f;
            if (alloc_dev(p, (void **)&ps.stage_tw_s, sizeof(float2) * Rst)) return MFFT_ERR_NOMEM;
            HIP_TRY(gen_twiddle(ps.stage_tw_s, Rst, Rst, 1, sign, 0, fac));
            if (ps.tw_N) {
                long long nlo = 1LL << ps.tw_bits;
                if (alloc_dev(p, (void **)&ps.tw_lo_s, sizeof(float2) * nlo)) return MFFT_ERR_NOMEM;
                HIP_TRY(gen_twiddle(ps.tw_lo_s, ps.tw_N, nlo, 1, sign, 0, fac));
            }
        }
        if (ps.onchip_cols) {
            // W_2048^e with the radix-16 stage scale, e in [0, 2048)
            if (alloc_dev(p, (void **)&ps.tw_lo_s, sizeof(float2) * kOnchipColsN)) return MFFT_ERR_NOMEM;
            HIP_TRY(gen_twiddle(ps.tw_lo_s, kOnchipColsN, kOnchipColsN, 1, sign, 0, ps.last_mul));
        }
        if (ps.small) ps.last_mul = p->d.scale == MFFT_SCALE_STAGE ? 1.0f / ps.R : 1.0f;
        if (ps.perm_axis >= 0 && p->d.in_order == MFFT_ORDER_DIGIT_REVERSED && !ps.rev_fast) {
            const std::vector<int> &rev = p->rev_axis[ps.perm_axis];
            if ((int)rev.size() != ps.R) return fail(MFFT_ERR_INVALID, "permutation size");
            if (alloc_dev(p, (void **)&ps.perm, sizeof(int) * ps.R)) return MFFT_ERR_NOMEM;
            HIP_TRY(hipMemcpy(ps.perm, rev.data(), sizeof(int) * ps.R, hipMemcpyHostToDevice));
        }
    }
    HIP_TRY(hipDeviceSynchronize());
    return MFFT_OK;
}

static int finalize_plan(mfft_plan_s *p)
{
    p->mid_fmt = (p->d.precision == MFFT_PREC_FP16) ? FMT_C16
                 : (p->d.precision == MFFT_PREC_BF16) ? FMT_CBF16
                                                      : FMT_C32;
    if (p->d.mid_fmt >= 0) p->mid_fmt = p->d.mid_fmt;
    int rc = finish_passes(p);
    if (rc) return rc;
    // scale bookkeeping
    double sc = 1.0;
    for (auto &ps : p->passes)
        for (int r : ps.radix) sc /= (p->d.scale == MFFT_SCALE_STAGE ? r : 1);
    p->out_scale = sc;
    // workspace and batch chunking
    bool need_work = false;
    for (auto &ps : p->passes)
        if (ps.src == BUF_WORK || ps.dst == BUF_WORK) need_work = true;
    long long wper = p->total;
    for (auto &ps : p->passes) wper = std::max(wper, ps.work_stride);
    long long per = wper * fmt_bytes(p->mid_fmt);
    long long limit = p->d.workspace_limit > 0 ? p->d.workspace_limit : (128LL << 20);
    p->chunk = p->batch;
    if (need_work) {
        long long c = std::max(1LL, limit / std::max(1LL, per));
        // as many chunks as the limit needs, of equal size (8 transforms under a limit of 7:
        // 4 + 4, not 7 + 1, whose one-transform launch would leave most of the device idle)
        const long long nch = (p->batch + c - 1) / c;
        p->chunk = std::min(p->batch, (p->batch + nch - 1) / nch);
        p->work_bytes = p->chunk * per;
        if (!p->d.external_workspace && alloc_dev(p, &p->work, (size_t)p->work_bytes)) return MFFT_ERR_NOMEM;
    }
    return build_tables(p);
}

static int check_desc(const mfft_desc *d)
{
    if (d->precision < 0 || d->precision > 3) return fail(MFFT_ERR_INVALID, "precision %d", d->precision);
    if (d->direction != -1 && d->direction != 1) return fail(MFFT_ERR_INVALID, "direction %d", d->direction);
    if (d->scale != 0 && d->scale != 1) return fail(MFFT_ERR_INVALID, "scale %d", d->scale);
    if (d->in_fmt < 0 || d->in_fmt > 2 || d->out_fmt < 0 || d->out_fmt > 2)
        return fail(MFFT_ERR_INVALID, "format");
    if (d->in_order != 0 && d->in_order != 1) return fail(MFFT_ERR_INVALID, "in_order");
    if (d->mid_fmt < -1 || d->mid_fmt > 2) return fail(MFFT_ERR_INVALID, "mid_fmt");
    if (d->split_hint < 0) return fail(MFFT_ERR_INVALID, "split_hint %d", d->split_hint);
    return MFFT_OK;
}

// ------------------------------------------------------------------ exec
static void *buf_ptr(int which, const void *in, void *out, void *work)
{
    return which == BUF_IN ? const_cast<void *>(in) : which == BUF_OUT ? out : work;
}

// Launch-trace record of one pass launch: kernel family, line length and formats, and the
// algorithmic bytes (every element read and written once; a fused real-image input reads
// the images' pixels, a fused multiplier its table once).
static void trace_pass(const mfft_plan_s *p, const PassSpec &ps, const TileArgs &a, long long nb, long long b0,
                       hipStream_t s)
{
    const char *fam = ps.small ? "stream"
                      : ps.fast_kind == KIND_ROWS ? "fast_rows"
                      : ps.fast_kind == KIND_ROWS_TR ? "fast_rows_tr"
                      : ps.fast_kind == KIND_COLS ? "fast_cols"
                      : ps.fast_kind == KIND_ONCHIP ? "onchip"
                      : ps.fast_kind == KIND_ONCHIP_COLS ? "onchip_cols"
                      : ps.generic ? "tile_generic" : "tile";
    const long long pts = (long long)a.c0 * a.c1 * a.c2 * ps.R;
    long long rd = pts * fmt_bytes(a.in_fmt);
    char src[16];
    snprintf(src, sizeof src, "%s", fmt_name(a.in_fmt));
    if (a.real_src) {
        // transforms b0 .. b0 + nb - 1: transform t holds channel t % C of images 2k and 2k + 1
        // (k = t / C; the second from real_src2 when given, absent past the batch)
        long long planes = 0;
        for (long long t = b0; t < b0 + nb; ++t) {
            const long long k = t / std::max(1, a.real_c);
            planes += (2 * k < a.real_b) + (a.real_src2 ? 1 : (2 * k + 1 < a.real_b));
        }
        rd = planes * a.real_h * a.real_w * (a.real_f64 ? 8 : 4);
        snprintf(src, sizeof src, "%s", a.real_f64 ? "real64" : "real32");
    }
    if (a.in_mul) rd += std::min(pts, a.in_mul_period) * 8;
    char name[64];
    snprintf(name, sizeof name, "%s R%d %s%s>%s", fam, ps.R, src, a.in_mul ? "*H" : "", fmt_name(a.out_fmt));
    (void)p;
    trace_launch(s, name, rd, pts * fmt_bytes(a.out_fmt));
}

// only >= 0: run that pass alone; kstart/kstop: events the runtime records around that
// pass's kernel (its own execution window; one launch per exec).
static int exec_chunk(mfft_plan_s *p, const char *in, char *out, long long nb, hipStream_t s, hipEvent_t *ev,
                      int &evi, void *work, const mfft::ExecFuse *fz = nullptr, long long b0 = 0, int only = -1,
                      hipEvent_t kstart = nullptr, hipEvent_t kstop = nullptr)
{
    int k = -1;
    for (auto &ps : p->passes) {
        if (++k != only && only >= 0) continue;
        TileArgs a;
        memset(&a, 0, sizeof a);
        if (fz && &ps == &p->passes.front()) {
            a.in_mul = fz->in_mul;
            a.in_mul_period = fz->in_mul_period;
            a.real_src = fz->real_src;
            a.real_src2 = fz->real_src2;
            a.real_f64 = fz->real_f64;
            a.real_b = fz->real_b;
            a.real_c = fz->real_c;
            a.real_h = fz->real_h;
            a.real_w = fz->real_w;
            a.real_pw = p->kind == 2 ? p->ny : p->n;
            a.real_p = p->total;
            a.batch_base = b0;
        }
        a.in = buf_ptr(ps.src, in, out, work);
        a.out = buf_ptr(ps.dst, in, out, work);
        a.in_fmt = src_fmt(p, ps);
        a.out_fmt = dst_fmt(p, ps);
        a.R = ps.R;
        a.nstages = (int)ps.radix.size();
        for (int i = 0; i < a.nstages; ++i) {
            a.radix[i] = ps.radix[i];
            a.stage_mul[i] = p->d.scale == MFFT_SCALE_STAGE ? 1.0f / ps.radix[i] : 1.0f;
        }
        a.T = ps.T;
        a.TL = ps.TL;
        auto cnt = [&](int li) -> long long {
            if (li < 0) return 1;
            return ps.lines[li].batch ? nb : ps.lines[li].count;
        };
        a.c0 = (int)cnt(ps.l0);
        a.c1 = (int)cnt(ps.l1);
        a.c2 = (int)cnt(ps.l2);
        a.in_s0 = ps.lines[ps.l0].in_stride;
        a.out_s0 = ps.lines[ps.l0].out_stride;
        a.in_s1 = ps.l1 >= 0 ? ps.lines[ps.l1].in_stride : 0;
        a.out_s1 = ps.l1 >= 0 ? ps.lines[ps.l1].out_stride : 0;
        a.in_s2 = ps.l2 >= 0 ? ps.lines[ps.l2].in_stride : 0;
        a.out_s2 = ps.l2 >= 0 ? ps.lines[ps.l2].out_stride : 0;
        a.in_ps = ps.ps_in;
        a.out_ps = ps.ps_out;
        a.in_perm = ps.perm;
        a.stage_tw = ps.stage_tw;
        a.dft_ops = p->ops;
        a.tw_N = ps.tw_N;
        a.tw_level = ps.tw_level == ps.l0 ? 0 : 1;
        a.tw_bits = ps.tw_bits;
        a.tw_lo = ps.tw_lo;
        a.tw_hi = ps.tw_hi;
        a.out_mul = 1.0f;
        a.sign = p->d.direction;
        a.in_vec = ((uintptr_t)a.in & 15) == 0;
        a.out_vec = ((uintptr_t)a.out & 15) == 0;
        a.stage_tw_s = ps.stage_tw_s;
        a.tw_lo_s = ps.tw_lo_s;
        a.last_mul = ps.last_mul;
        a.tail_mul = ps.tail_mul;
        a.rev_input = ps.rev_fast ? 1 : 0;
        a.in_rev = ps.rev_q;
        a.tw_rev = ps.rev_tw;
        memcpy(a.w16, g_w16, sizeof a.w16);
        long long ntiles = ((a.c0 + a.T - 1) / a.T) * (long long)a.c1 * a.c2;
        if (ntiles <= 0) continue;
        if (ntiles >= (1LL << 31)) return fail(MFFT_ERR_UNSUPPORTED, "grid too large");
        if (trace_active()) trace_pass(p, ps, a, nb, b0, s);
        if (ev) HIP_TRY(hipEventRecord(ev[evi++], s));
        auto launch = [&](dim3 grid, dim3 block) -> hipError_t {
            void *args[] = {&a};
            if (kstart || kstop) return hipExtLaunchKernel(ps.fn, grid, block, args, 0, s, kstart, kstop, 0);
            return hipLaunchKernel(ps.fn, grid, block, args, 0, s);
        };
        if (ps.fast_kind == KIND_ONCHIP_COLS) {
            // one workgroup per 32 adjacent columns of one image
            OnchipColsArgs oa;
            memset(&oa, 0, sizeof oa);
            oa.in = a.in;
            oa.out = a.out;
            oa.col_blocks = (int)(ps.lines[0].count / kOnchipColsLines);
            oa.blocks = (long long)oa.col_blocks * nb;
            oa.ps = (int)ps.ps_in;
            oa.ps_out = (int)ps.ps_out;
            oa.in_dist = ps.lines[1].in_stride;
            oa.out_dist = ps.lines[1].out_stride;
            oa.dft_ops = p->ops;
            oa.tw_s = ps.tw_lo_s;
            oa.stw_s = ps.stage_tw_s;
            oa.tail_mul = ps.tail_mul;
            oa.sign = p->d.direction;
            memcpy(oa.w16, g_w16, sizeof oa.w16);
            void *oargs[] = {&oa};
            if (oa.blocks >= (1LL << 31)) return fail(MFFT_ERR_UNSUPPORTED, "grid too large");
            const dim3 grid((unsigned)oa.blocks), block(kOnchipThreads);
            if (kstart || kstop)
                HIP_TRY(hipExtLaunchKernel(ps.fn, grid, block, oargs, 0, s, kstart, kstop, 0));
            else
                HIP_TRY(hipLaunchKernel(ps.fn, grid, block, oargs, 0, s));
            continue;
        }
        if (ps.fast_kind == KIND_ONCHIP) {
            // one workgroup per line; a digit-reversed line is read in 16-byte pieces
            if (!ps.rev_list.empty() && ((uintptr_t)a.in & 15))
                return fail(MFFT_ERR_INVALID, "digit-reversed 65536-point input must be 16-byte aligned");
            OnchipArgs oa;
            oa.in = a.in;
            oa.out = a.out;
            oa.batch = nb;
            oa.in_dist = ps.lines[0].in_stride;
            oa.out_dist = ps.lines[0].out_stride;
            oa.dft_ops = p->ops;
            oa.stage_tw_s = ps.stage_tw_s;
            oa.tw_lo_s = ps.tw_lo_s;
            oa.tw_hi = ps.tw_hi;
            oa.fac = ps.last_mul;
            void *oargs[] = {&oa};
            const dim3 grid((unsigned)nb), block(kOnchipThreads);
            if (kstart || kstop)
                HIP_TRY(hipExtLaunchKernel(ps.fn, grid, block, oargs, 0, s, kstart, kstop, 0));
            else
                HIP_TRY(hipLaunchKernel(ps.fn, grid, block, oargs, 0, s));
            continue;
        }
        if (ps.small) {
            long long nl = (long long)a.c0 * a.c1 * a.c2;
            if (a.in_mul) nl = in_mul_threads((long long)a.c0 * a.c1, a.c2);
            // largest byte offset a line's point reaches, per side (strides are >= 0)
            auto span = [&](long long s0, long long s1, long long s2, long long sp, int fmt) -> long long {
                if (s0 < 0 || s1 < 0 || s2 < 0 || sp < 0) return -1;
                return ((a.c0 - 1) * s0 + (a.c1 - 1) * s1 + (a.c2 - 1) * s2 + (a.R - 1) * sp) * fmt_bytes(fmt);
            };
            const long long si = span(a.in_s0, a.in_s1, a.in_s2, a.in_ps, a.in_fmt);
            const long long so = span(a.out_s0, a.out_s1, a.out_s2, a.out_ps, a.out_fmt);
            a.off32 = (si >= 0 && si < (1LL << 32) - 8 ? 1 : 0) | (so >= 0 && so < (1LL << 32) - 8 ? 2 : 0);
            a.in_span = (a.off32 & 1) ? (unsigned)(si + fmt_bytes(a.in_fmt)) : 0u;
            a.out_span = (a.off32 & 2) ? (unsigned)(so + fmt_bytes(a.out_fmt)) : 0u;
            HIP_TRY(launch(dim3((unsigned)((nl + 255) / 256)), dim3(256)));
            continue;
        }
        if (ps.fast_kind >= 0) {
            const int twc = ps.fast_kind == KIND_COLS ? kFastTwCols : kFastTw;
            a.tw_lds = ps.tw_N && (1 << ps.tw_bits) <= twc && ((ps.tw_N + (1 << ps.tw_bits) - 1) >> ps.tw_bits) <= twc;
            long long grid = std::min(ntiles, ps.max_grid);
            if (grid > 8) grid -= grid % 8;
            HIP_TRY(launch(dim3((unsigned)grid), dim3(fast_threads(p->d.precision, ps.R, ps.fast_kind))));
        } else {
            HIP_TRY(launch_tile(ps.fn, a, ntiles, s, kstart, kstop));
        }
    }
    return MFFT_OK;
}

// ------------------------------------------------------------------ C ABI
namespace mfft {
void mfft_trace_reset(hipEvent_t *events, mfft_launch_record *rec, int cap);
hipError_t mfft_trace_stop(int *n, bool *overflow, hipEvent_t *last);
}  // namespace mfft

extern "C" {

int mfft_desc_default(mfft_desc *d)
{
    if (!d) return fail(MFFT_ERR_INVALID, "null desc");
    memset(d, 0, sizeof *d);
    d->precision = MFFT_PREC_SPLIT_FP16;
    d->direction = MFFT_FORWARD;
    d->scale = MFFT_SCALE_STAGE;
    d->in_fmt = MFFT_FMT_C16;
    d->out_fmt = MFFT_FMT_C16;
    d->in_order = MFFT_ORDER_NATURAL;
    d->mid_fmt = -1;
    return MFFT_OK;
}

// Plan creation is the set-up of mfft_exec_noalloc too: the address of its workspace array
// on the current device is resolved here, so an exec under graph capture or in a serving
// loop finds it cached (a failure is reported by that exec, not here).
static void resolve_noalloc_scratch()
{
    const std::string keep = g_err;
    void *ws = nullptr;
    (void)noalloc_scratch(&ws);
    g_err = keep;
}

// split_hint = MFFT_SPLIT_TUNED: take split and chunk from the autotuner's cache.
static mfft_desc resolve_tuned(const mfft_desc *d, int nx, int ny, int64_t batch)
{
    mfft_desc r = *d;
    if (d->split_hint != MFFT_SPLIT_TUNED) return r;
    r.split_hint = 0;
    mfft_tune_entry e;
    std::string keep = g_err;
    if (mfft_tune_lookup(nullptr, nx, ny, batch, d, &e) == MFFT_OK) {
        r.split_hint = e.split_hint;
        r.workspace_limit = e.workspace_limit;
    }
    g_err = keep;
    return r;
}

static int create_1d(mfft_plan *plan, int n, int64_t batch, const mfft_desc *d_in, bool allow_onchip)
{
    if (!plan || !d_in) return fail(MFFT_ERR_INVALID, "null argument");
    *plan = nullptr;
    if (n < 2 || batch < 1) return fail(MFFT_ERR_INVALID, "n=%d batch=%lld", n, (long long)batch);
    const mfft_desc dt = resolve_tuned(d_in, n, 0, batch);
    const mfft_desc *d = &dt;
    int rc = check_desc(d);
    if (rc) return rc;
    mfft_plan_s *p = new mfft_plan_s;
    p->kind = 1;
    p->n = n;
    p->batch = batch;
    p->d = *d;
    p->total = n;
    const bool rev = d->in_order == MFFT_ORDER_DIGIT_REVERSED;
    if (rev) {
        rev_radices(n, p->rev_list_axis[0]);
        gen_rev(n, p->rev_list_axis[0], p->rev_axis[0]);
    }
    if (allow_onchip && batch >= onchip_min_lines() && plan_onchip(p, d)) {
        rc = MFFT_OK;
    } else if (allow_onchip && n >= (1 << 19) && n <= (1 << 20) &&
               batch >= onchip_rows_min_lines(n / kOnchipColsN) && plan_onchip_rows(p, d)) {
        rc = MFFT_OK;
    } else if (rev && n > kMaxTile) {
        rc = plan_reversed_1d(p);
    } else {
        std::vector<Level> lines{Level{batch, n, n, true}};
        rc = plan_axis(p->passes, n, 1, 1, lines, BUF_IN, BUF_OUT, rev ? 0 : -1, d->split_hint, narrow_mid(d));
    }
    if (!rc) rc = finalize_plan(p);
    if (rc) {
        std::string msg = g_err;
        mfft_plan_destroy(p);
        g_err = msg;
        return rc;
    }
    resolve_noalloc_scratch();
    *plan = p;
    return MFFT_OK;
}

static bool has_onchip(const mfft_plan_s *p)
{
    for (auto &ps : p->passes)
        if (ps.onchip || ps.onchip_cols) return true;
    return false;
}

int mfft_plan_create_1d(mfft_plan *plan, int n, int64_t batch, const mfft_desc *d_in)
{
    int rc = create_1d(plan, n, batch, d_in, true);
    if (rc || !has_onchip(*plan)) return rc;
    mfft_plan p = *plan, q = nullptr;
    const long long below = n == kOnchipN ? onchip_min_lines() : onchip_rows_min_lines(n / kOnchipColsN);
    rc = create_1d(&q, n, below - 1, d_in, false);
    if (rc) {
        std::string msg = g_err;
        mfft_plan_destroy(p);
        *plan = nullptr;
        g_err = msg;
        return rc;
    }
    p->small = q;
    p->small_below = below;
    return MFFT_OK;
}

static int create_2d(mfft_plan *plan, int nx, int ny, int64_t batch, const mfft_desc *d_in, bool allow_onchip,
                     bool out_mid = false)
{
    if (!plan || !d_in) return fail(MFFT_ERR_INVALID, "null argument");
    *plan = nullptr;
    if (nx < 2 || ny < 2 || batch < 1) return fail(MFFT_ERR_INVALID, "nx=%d ny=%d batch=%lld", nx, ny, (long long)batch);
    const mfft_desc dt = resolve_tuned(d_in, nx, ny, batch);
    const mfft_desc *d = &dt;
    int rc = check_desc(d);
    if (rc) return rc;
    mfft_plan_s *p = new mfft_plan_s;
    p->kind = 2;
    p->nx = nx;
    p->ny = ny;
    p->n = nx * ny;
    p->batch = batch;
    p->d = *d;
    p->total = (long long)nx * ny;
    bool rev = d->in_order == MFFT_ORDER_DIGIT_REVERSED;
    if (rev) {
        rev_radices(nx, p->rev_list_axis[0]);
        gen_rev(nx, p->rev_list_axis[0], p->rev_axis[0]);
        rev_radices(ny, p->rev_list_axis[1]);
        gen_rev(ny, p->rev_list_axis[1], p->rev_axis[1]);
    }
    // the column pass reads the x order itself when its kernel is the strided matrix-core
    // one (then up to 1024 points stay one pass); its input is the rows' intermediate
    DigitRev qx;
    const int col_in_fmt = d->mid_fmt >= 0                          ? d->mid_fmt
                           : d->precision == MFFT_PREC_FP16 ? MFFT_FMT_C16
                           : d->precision == MFFT_PREC_BF16 ? MFFT_FMT_CBF16
                                                            : MFFT_FMT_C32;
    const bool rev_cols = rev && fast_rev_ok(d->precision, col_in_fmt) && fast_rev_spec(nx, p->rev_list_axis[0], qx);
    long long NN = (long long)nx * ny;
    // rows (y axis, contiguous): lines = (x, b)
    std::vector<Level> rows{Level{nx, ny, ny, false}, Level{batch, NN, NN, true}};
    const int odd_hint = d->split_hint == 1 ? 1 : 0;  // 1 keeps odd * 2^a axes in one pass
    rc = plan_axis(p->passes, ny, 1, 1, rows, BUF_IN, BUF_WORK, rev ? 1 : -1, odd_hint, narrow_mid(d));
    // columns (x axis, stride ny): lines = (y, b)
    if (!rc) {
        // split_hint > 1: the first-pass length of the column axis when it is split
        const int col_hint = d->split_hint > 1 ? d->split_hint : odd_hint;
        std::vector<Level> cols{Level{ny, 1, 1, false}, Level{batch, NN, NN, true}};
        // 2048-point columns with 16-bit words on both sides, a multiple of 32 of them: one
        // pass of the on-chip column kernel (the split into a strided and a streaming pass
        // moves every point twice)
        const int fmt16 = d->precision == MFFT_PREC_BF16 ? MFFT_FMT_CBF16 : MFFT_FMT_C16;
        if (allow_onchip && nx == kOnchipColsN && ny % kOnchipColsLines == 0 && batch >= onchip_cols_min_images(ny) &&
            !rev && d->split_hint <= 1 && col_in_fmt == fmt16 && d->out_fmt == fmt16 &&
            find_onchip_cols_kernel(d->precision, fmt16)) {
            PassSpec ps;
            ps.R = nx;
            ps.ps_in = ps.ps_out = ny;
            ps.lines = cols;
            ps.src = BUF_WORK;
            ps.dst = BUF_OUT;
            ps.onchip_cols = true;
            p->passes.push_back(ps);
        } else {
            rc = plan_axis(p->passes, nx, ny, ny, cols, BUF_WORK, BUF_OUT, rev ? 0 : -1, col_hint, narrow_mid(d),
                           rev_cols);
        }
    }
    p->out_mid = out_mid;
    if (!rc) rc = finalize_plan(p);
    if (rc) {
        std::string msg = g_err;
        mfft_plan_destroy(p);
        g_err = msg;
        return rc;
    }
    resolve_noalloc_scratch();
    *plan = p;
    return MFFT_OK;
}

static int create_2d_all(mfft_plan *plan, int nx, int ny, int64_t batch, const mfft_desc *d_in, bool out_mid)
{
    int rc = create_2d(plan, nx, ny, batch, d_in, true, out_mid);
    if (rc || !has_onchip(*plan)) return rc;
    mfft_plan p = *plan, q = nullptr;
    const long long below = onchip_cols_min_images(ny);
    rc = create_2d(&q, nx, ny, below - 1, d_in, false, out_mid);
    if (rc) {
        std::string msg = g_err;
        mfft_plan_destroy(p);
        *plan = nullptr;
        g_err = msg;
        return rc;
    }
    p->small = q;
    p->small_below = below;
    return MFFT_OK;
}

int mfft_plan_create_2d(mfft_plan *plan, int nx, int ny, int64_t batch, const mfft_desc *d_in)
{
    return create_2d_all(plan, nx, ny, batch, d_in, false);
}

int mfft_plan_destroy(mfft_plan plan)
{
    if (!plan) return MFFT_OK;
    (void)hipDeviceSynchronize();
    if (plan->small) mfft_plan_destroy(plan->small);
    for (void *ptr : plan->allocs) (void)hipFree(ptr);
    plan->ws_order.destroy();
    delete plan;
    return MFFT_OK;
}

int mfft_plan_info(mfft_plan p, mfft_info *info)
{
    if (!p || !info) return fail(MFFT_ERR_INVALID, "null argument");
    memset(info, 0, sizeof *info);
    info->kind = p->kind;
    info->n = p->n;
    info->nx = p->nx;
    info->ny = p->ny;
    info->batch = p->batch;
    info->npasses = (int)std::min<size_t>(p->passes.size(), MFFT_MAX_PASSES);
    long long nchunks = (p->batch + p->chunk - 1) / p->chunk;
    for (int i = 0; i < info->npasses; ++i) {
        const PassSpec &ps = p->passes[i];
        mfft_pass_info &pi = info->pass[i];
        pi.R = ps.R;
        pi.lines_per_tile = ps.T;
        pi.lane_lines = ps.TL;
        pi.nstages = (int)ps.radix.size();
        for (int k = 0; k < pi.nstages && k < 8; ++k) pi.radix[k] = ps.radix[k];
        long long lines = 1;
        for (auto &l : ps.lines) lines *= l.batch ? p->chunk : l.count;
        long long c0 = ps.lines[ps.l0].batch ? p->chunk : ps.lines[ps.l0].count;
        pi.tiles = (lines / c0) * ((c0 + ps.T - 1) / ps.T);
        pi.generic = ps.generic;
        pi.in_elem_bytes = fmt_bytes(src_fmt(p, ps));
        pi.out_elem_bytes = fmt_bytes(dst_fmt(p, ps));
        pi.points = lines * ps.R;
        pi.kernel = ps.small ? MFFT_KERNEL_STREAM
                    : ps.fast_kind == KIND_ROWS ? MFFT_KERNEL_ROWS
                    : ps.fast_kind == KIND_ROWS_TR ? MFFT_KERNEL_ROWS_TR
                    : ps.fast_kind == KIND_COLS ? MFFT_KERNEL_COLS
                    : ps.fast_kind == KIND_ONCHIP ? MFFT_KERNEL_ONCHIP
                    : ps.fast_kind == KIND_ONCHIP_COLS ? MFFT_KERNEL_ONCHIP_COLS
                    : ps.generic ? MFFT_KERNEL_GENERIC : MFFT_KERNEL_TILE;
    }
    info->out_scale = p->out_scale;
    // a caller-supplied workspace also serves the companion plan of small execs
    info->workspace_bytes = std::max(p->work_bytes, p->small ? p->small->work_bytes : 0LL);
    info->batch_chunk = p->chunk;
    info->launches_per_exec = nchunks * (long long)p->passes.size();
    return MFFT_OK;
}

static int exec_all(mfft_plan p, const void *in, void *out, void *stream, hipEvent_t *ev, long long batch = -1,
                    void *work = nullptr, const mfft::ExecFuse *fz = nullptr, int only = -1,
                    hipEvent_t kstart = nullptr, hipEvent_t kstop = nullptr)
{
    if (p->small && batch >= 0 && batch < p->small_below && only < 0 && !ev &&
        (!fz || (p->small->passes.front().small && p->small->passes.front().src == BUF_IN)))
        return exec_all(p->small, in, out, stream, nullptr, batch, work, fz, -1, nullptr, nullptr);
    if (p->out_mid && in == out)
        for (const PassSpec &ps : p->passes) {
            // a pass that reads the input (not a fused real source) and writes the output in
            // another format would overwrite lines it has not read yet
            const bool reads = !(fz && fz->real_src && &ps == &p->passes.front());
            if (reads && ps.src == BUF_IN && ps.dst == BUF_OUT && src_fmt(p, ps) != dst_fmt(p, ps))
                return fail(MFFT_ERR_INVALID, "plan keeps intermediates in the output buffer: in and out must differ");
        }
    std::unique_lock<std::mutex> lk(p->exec_mu, std::defer_lock);
    if (!work) lk.lock();  // the plan's own workspace is shared by its callers
    hipStream_t s = (hipStream_t)stream;
    long long ib = p->total * fmt_bytes(p->d.in_fmt), ob = p->total * fmt_bytes(p->d.out_fmt);
    if (batch < 0) batch = p->batch;
    if (!work && !p->work && p->work_bytes > 0) return fail(MFFT_ERR_INVALID, "plan needs a workspace (mfft_exec_ex)");
    // the plan's own workspace: an exec on another stream than the previous one runs
    // after it (two streams sharing p->work would otherwise overwrite each other's
    // intermediates between the passes)
    const bool own_ws = !work && p->work;
    if (own_ws) HIP_TRY(p->ws_order.acquire(s));
    int evi = 0, rc = MFFT_OK;
    for (long long b0 = 0; b0 < batch && !rc; b0 += p->chunk) {
        long long nb = std::min(p->chunk, batch - b0);
        rc = exec_chunk(p, (const char *)in + b0 * ib, (char *)out + b0 * ob, nb, s, ev, evi, work ? work : p->work, fz,
                        b0, only, kstart, kstop);
    }
    if (own_ws) {
        std::string keep = g_err;
        hipError_t e = p->ws_order.release(s);  // also after a partial enqueue
        if (!rc && e != hipSuccess) return fail(MFFT_ERR_HIP, "workspace order: %s", hipGetErrorString(e));
        if (rc) g_err = keep;
    }
    if (rc) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[evi++], s));
    return MFFT_OK;
}

int mfft_exec(mfft_plan p, const void *in, void *out, void *stream)
{
    if (!p || !in || !out) return fail(MFFT_ERR_INVALID, "null argument");
    return exec_all(p, in, out, stream, nullptr);
}

int mfft_exec_ex(mfft_plan p, const void *in, void *out, int64_t batch, void *workspace, void *stream)
{
    if (!p || !in || !out) return fail(MFFT_ERR_INVALID, "null argument");
    if (batch < 1 || batch > p->batch) return fail(MFFT_ERR_INVALID, "batch %lld outside 1..%lld", (long long)batch, p->batch);
    return exec_all(p, in, out, stream, nullptr, batch, workspace);
}

// ------------------------------------------------------------------ work-queue batcher
struct mfft_batcher_s {
    mfft_desc d{};
    int max_items = 0;
    DftOps *ops = nullptr;
    std::vector<std::pair<int, float2 *>> tables;  // stage twiddles per length
    struct Slot {
        TileArgs *h_jobs = nullptr, *d_jobs = nullptr;
        long long *h_first = nullptr, *d_first = nullptr;
        unsigned long long *counter = nullptr;
        hipEvent_t done = nullptr;
        bool used = false;
    } slot[4];
    int next = 0;
    int grid = 0;
    std::mutex mu;
};

int mfft_batcher_destroy(mfft_batcher b)
{
    if (!b) return MFFT_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(b->ops);
    for (auto &t : b->tables) (void)hipFree(t.second);
    for (auto &sl : b->slot) {
        (void)hipHostFree(sl.h_jobs);
        (void)hipHostFree(sl.h_first);
        (void)hipFree(sl.d_jobs);
        (void)hipFree(sl.d_first);
        (void)hipFree(sl.counter);
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
    delete b;
    return MFFT_OK;
}

int mfft_batcher_create(mfft_batcher *out, const mfft_desc *d, int max_items)
{
    if (!out || !d || max_items < 1) return fail(MFFT_ERR_INVALID, "bad argument");
    int rc = check_desc(d);
    if (rc) return rc;
    if (ensure_w16()) return MFFT_ERR_HIP;
    mfft_batcher b = new mfft_batcher_s;
    b->d = *d;
    b->max_items = max_items;
    hipError_t e = hipMalloc((void **)&b->ops, sizeof(DftOps));
    if (e == hipSuccess) e = gen_dft_ops(b->ops, d->direction, 0);
    for (auto &sl : b->slot) {
        if (e == hipSuccess) e = hipHostMalloc((void **)&sl.h_jobs, sizeof(TileArgs) * max_items, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void **)&sl.h_first, sizeof(long long) * (max_items + 1), hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc((void **)&sl.d_jobs, sizeof(TileArgs) * max_items);
        if (e == hipSuccess) e = hipMalloc((void **)&sl.d_first, sizeof(long long) * (max_items + 1));
        if (e == hipSuccess) e = hipMalloc((void **)&sl.counter, sizeof(unsigned long long));
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        mfft_batcher_destroy(b);
        return fail(MFFT_ERR_HIP, "batcher: %s", hipGetErrorString(e));
    }
    b->grid = queue_grid(d->precision);
    *out = b;
    return MFFT_OK;
}

static float2 *batch_table(mfft_batcher b, int n, hipStream_t s)
{
    for (auto &t : b->tables)
        if (t.first == n) return t.second;
    float2 *tab = nullptr;
    if (hipMalloc((void **)&tab, sizeof(float2) * n) != hipSuccess) return nullptr;
    // generated once on the null stream and waited for, so any stream may use it after
    (void)s;
    if (gen_twiddle(tab, n, n, 1, b->d.direction, 0) != hipSuccess) return nullptr;
    if (hipStreamSynchronize(0) != hipSuccess) return nullptr;
    b->tables.push_back({n, tab});
    return tab;
}

int mfft_batcher_exec(mfft_batcher b, const mfft_batch_item *items, int nitems, void *stream)
{
    if (!b || (!items && nitems)) return fail(MFFT_ERR_INVALID, "null argument");
    if (nitems < 0 || nitems > b->max_items) return fail(MFFT_ERR_INVALID, "%d items (max %d)", nitems, b->max_items);
    if (nitems == 0) return MFFT_OK;
    std::lock_guard<std::mutex> lk(b->mu);
    hipStream_t s = (hipStream_t)stream;
    auto &sl = b->slot[b->next];
    b->next = (b->next + 1) % 4;
    if (sl.used) HIP_TRY(hipEventSynchronize(sl.done));  // the slot's previous launch has read its jobs
    long long total = 0;
    for (int i = 0; i < nitems; ++i) {
        const mfft_batch_item &it = items[i];
        std::vector<int> rad;
        if (it.n < 2 || it.n > kMaxTile || !line_radices(it.n, rad) || it.count < 1 || !it.in || !it.out)
            return fail(MFFT_ERR_INVALID, "item %d: n=%d count=%lld", i, it.n, (long long)it.count);
        float2 *tab = batch_table(b, it.n, s);
        if (!tab) return fail(MFFT_ERR_NOMEM, "twiddle table");
        TileArgs &a = sl.h_jobs[i];
        memset(&a, 0, sizeof a);
        a.in = it.in;
        a.out = it.out;
        a.in_fmt = b->d.in_fmt;
        a.out_fmt = b->d.out_fmt;
        a.R = it.n;
        a.nstages = (int)rad.size();
        for (int k = 0; k < a.nstages; ++k) {
            a.radix[k] = rad[k];
            a.stage_mul[k] = b->d.scale == MFFT_SCALE_STAGE ? 1.0f / rad[k] : 1.0f;
        }
        a.T = std::max(1, kMaxTile / it.n);
        a.TL = 1;
        a.c0 = (int)it.count;
        a.c1 = a.c2 = 1;
        a.in_s0 = a.out_s0 = it.n;
        a.in_ps = a.out_ps = 1;
        a.stage_tw = tab;
        a.dft_ops = b->ops;
        a.out_mul = 1.0f;
        a.sign = b->d.direction;
        a.in_vec = ((uintptr_t)it.in & 15) == 0;
        a.out_vec = ((uintptr_t)it.out & 15) == 0;
        memcpy(a.w16, g_w16, sizeof a.w16);
        sl.h_first[i] = total;
        total += (it.count + a.T - 1) / a.T;
    }
    sl.h_first[nitems] = total;
    HIP_TRY(hipMemcpyAsync(sl.d_jobs, sl.h_jobs, sizeof(TileArgs) * nitems, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sl.d_first, sl.h_first, sizeof(long long) * (nitems + 1), hipMemcpyHostToDevice, s));
    int grid = (int)std::min<long long>(b->grid, total);
    if (trace_active()) {
        // one record for the counter reset and the work-queue kernel
        long long rd = 0, wr = 0;
        for (int i = 0; i < nitems; ++i) {
            rd += items[i].count * items[i].n * (long long)fmt_bytes(b->d.in_fmt);
            wr += items[i].count * items[i].n * (long long)fmt_bytes(b->d.out_fmt);
        }
        char name[64];
        snprintf(name, sizeof name, "queue %d items %s>%s", nitems, fmt_name(b->d.in_fmt), fmt_name(b->d.out_fmt));
        trace_launch(s, name, rd, wr);
    }
    HIP_TRY(launch_queue(b->d.precision, sl.d_jobs, sl.d_first, nitems, total, sl.counter, grid, s));
    HIP_TRY(hipEventRecord(sl.done, s));
    sl.used = true;
    return MFFT_OK;
}

int mfft_exec_pass(mfft_plan p, int pass, const void *in, void *out, void *stream, void *start_event, void *stop_event)
{
    if (!p || !in || !out) return fail(MFFT_ERR_INVALID, "null argument");
    if (pass < 0 || pass >= (int)p->passes.size())
        return fail(MFFT_ERR_INVALID, "pass %d of %d", pass, (int)p->passes.size());
    if ((start_event || stop_event) && p->chunk < p->batch)
        return fail(MFFT_ERR_UNSUPPORTED, "kernel events need a single-chunk plan (%lld per chunk, batch %lld)",
                    p->chunk, p->batch);
    return exec_all(p, in, out, stream, nullptr, -1, nullptr, nullptr, pass, (hipEvent_t)start_event,
                    (hipEvent_t)stop_event);
}

int mfft_exec_events(mfft_plan p, const void *in, void *out, void *stream, void **events, int nevents)
{
    if (!p || !in || !out || !events) return fail(MFFT_ERR_INVALID, "null argument");
    long long nchunks = (p->batch + p->chunk - 1) / p->chunk;
    long long need = nchunks * (long long)p->passes.size() + 1;
    if (nevents < need) return fail(MFFT_ERR_INVALID, "need %lld events, got %d", need, nevents);
    return exec_all(p, in, out, stream, (hipEvent_t *)events);
}

struct mfft_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int mfft_graph_create(mfft_graph *graph, mfft_plan p, const void *in, void *out, void *stream)
{
    if (!graph || !p || !in || !out) return fail(MFFT_ERR_INVALID, "null argument");
    *graph = nullptr;
    hipStream_t s = (hipStream_t)stream;
    hipStream_t cap = s;
    bool own = false;
    if (!cap) {  // the legacy default stream cannot be captured
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        own = true;
    }
    mfft_graph g = new mfft_graph_s;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    int rc = e == hipSuccess ? exec_all(p, in, out, cap, nullptr) : fail(MFFT_ERR_HIP, "begin capture: %s", hipGetErrorString(e));
    hipError_t e2 = hipStreamEndCapture(cap, &g->graph);
    if (!rc && e2 != hipSuccess) rc = fail(MFFT_ERR_HIP, "end capture: %s", hipGetErrorString(e2));
    if (!rc && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        rc = fail(MFFT_ERR_HIP, "graph instantiate");
    if (own) (void)hipStreamDestroy(cap);
    if (rc) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return rc;
    }
    *graph = g;
    return MFFT_OK;
}

int mfft_graph_launch(mfft_graph g, void *stream)
{
    if (!g) return fail(MFFT_ERR_INVALID, "null graph");
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return MFFT_OK;
}

int mfft_graph_destroy(mfft_graph g)
{
    if (!g) return MFFT_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return MFFT_OK;
}

int mfft_input_permutation(mfft_plan p, int axis, int *rev, int n)
{
    if (!p || !rev || axis < 0 || axis > 1) return fail(MFFT_ERR_INVALID, "bad argument");
    int len = p->kind == 1 ? p->n : (axis == 0 ? p->nx : p->ny);
    if (n != len) return fail(MFFT_ERR_INVALID, "length %d != %d", n, len);
    std::vector<int> rad, r;
    rev_radices(len, rad);
    gen_rev(len, rad, r);
    memcpy(rev, r.data(), sizeof(int) * len);
    return MFFT_OK;
}

int mfft_stage_convert(const void *src, int src_fmt, void *dst, int dst_fmt, int64_t n_inner, int64_t batch,
                       const int *perm_x, int nx, const int *perm_y, int ny, double scale, void *stream)
{
    if (!src || !dst || n_inner < 0 || batch < 0) return fail(MFFT_ERR_INVALID, "bad argument");
    if (src_fmt < 0 || src_fmt > 3 || dst_fmt < 0 || dst_fmt > 3) return fail(MFFT_ERR_INVALID, "format");
    if (perm_y && (ny <= 0 || n_inner % ny != 0)) return fail(MFFT_ERR_INVALID, "2D staging needs ny | n_inner");
    if (src == dst) return fail(MFFT_ERR_INVALID, "staging is out of place");
    if (trace_active()) {
        auto eb = [](int f) -> long long { return f == MFFT_FMT_C64 ? 16 : f == MFFT_FMT_C32 ? 8 : 4; };
        trace_launch((hipStream_t)stream, "stage_convert", batch * n_inner * eb(src_fmt), batch * n_inner * eb(dst_fmt));
    }
    HIP_TRY(stage_convert(src, src_fmt, dst, dst_fmt, n_inner, batch, perm_x, nx, perm_y, ny, scale,
                          (hipStream_t)stream));
    return MFFT_OK;
}

int mfft_trace_begin(void **events, mfft_launch_record *records, int capacity)
{
    if (!events || !records || capacity < 1) return fail(MFFT_ERR_INVALID, "trace: bad argument");
    if (mfft::trace_active()) return fail(MFFT_ERR_INVALID, "trace: already active on this thread");
    mfft_trace_reset((hipEvent_t *)events, records, capacity);
    return MFFT_OK;
}

int mfft_trace_end(void *stream, int *nlaunches)
{
    if (!nlaunches) return fail(MFFT_ERR_INVALID, "trace: null argument");
    int n = 0;
    bool overflow = false;
    hipEvent_t last = nullptr;
    hipError_t err = mfft_trace_stop(&n, &overflow, &last);
    *nlaunches = n;
    if (!last) return fail(MFFT_ERR_INVALID, "trace: not active on this thread");
    if (err != hipSuccess) return fail(MFFT_ERR_HIP, "trace: event record: %s", hipGetErrorString(err));
    HIP_TRY(hipEventRecord(last, (hipStream_t)stream));
    if (overflow) return fail(MFFT_ERR_INVALID, "trace: more launches than the %d records", n);
    return MFFT_OK;
}

const char *mfft_last_error(void) { return g_err.c_str(); }

const char *mfft_version(void) { return "mfft 0.1 (gfx950 matrix-core FFT)"; }

}  // extern "C"

namespace mfft {

void set_last_error(const char *msg) { g_err = msg ? msg : ""; }

void digit_reversal_radices(int n, std::vector<int> &out) { rev_radices(n, out); }

// ------------------------------------------------------------------ launch trace
namespace {
struct Trace {
    hipEvent_t *events = nullptr;
    mfft_launch_record *rec = nullptr;
    int cap = 0, n = 0;
    bool overflow = false;
    hipError_t err = hipSuccess;
};
thread_local Trace t_trace;
thread_local bool t_trace_on = false;
}  // namespace

bool trace_active() { return t_trace_on; }

void mfft_trace_reset(hipEvent_t *events, mfft_launch_record *rec, int cap)
{
    t_trace = Trace();
    t_trace.events = events;
    t_trace.rec = rec;
    t_trace.cap = cap;
    t_trace_on = true;
}

// ends the trace: n launches recorded; *last = the event to record after them (null when no
// trace was active)
hipError_t mfft_trace_stop(int *n, bool *overflow, hipEvent_t *last)
{
    *last = nullptr;
    if (!t_trace_on) return hipSuccess;
    t_trace_on = false;
    *n = t_trace.n;
    *overflow = t_trace.overflow;
    *last = t_trace.events[t_trace.n];
    return t_trace.err;
}

const char *fmt_name(int fmt) { return fmt == FMT_C16 ? "c16" : fmt == FMT_CBF16 ? "cbf16" : "c32"; }

void trace_launch(hipStream_t s, const char *kernel, long long bytes_read, long long bytes_written)
{
    if (!t_trace_on) return;
    Trace &t = t_trace;
    if (t.n >= t.cap) {
        t.overflow = true;
        return;
    }
    const hipError_t e = hipEventRecord(t.events[t.n], s);
    if (e != hipSuccess && t.err == hipSuccess) t.err = e;
    mfft_launch_record &r = t.rec[t.n];
    snprintf(r.kernel, sizeof r.kernel, "%s", kernel);
    r.bytes_read = bytes_read;
    r.bytes_written = bytes_written;
    ++t.n;
}

static bool capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

hipError_t StreamOrder::acquire(hipStream_t s)
{
    if (!used || s == last || capturing(s)) return hipSuccess;
    if (ev) return hipStreamWaitEvent(s, ev, 0);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    return e;
}

hipError_t StreamOrder::release(hipStream_t s)
{
    if (capturing(s)) return hipSuccess;
    last = s;
    used = true;
    return ev ? hipEventRecord(ev, s) : hipSuccess;
}

void StreamOrder::destroy()
{
    if (ev) (void)hipEventDestroy(ev);
    ev = nullptr;
}

int plan_create_2d_out_mid(mfft_plan *plan, int nx, int ny, int64_t batch, const mfft_desc *d)
{
    return create_2d_all(plan, nx, ny, batch, d, true);
}

bool plan_first_pass_fusable(mfft_plan p)
{
    return p && !p->passes.empty() && p->passes.front().small && p->passes.front().src == BUF_IN;
}

int exec_fused(mfft_plan p, const void *in, void *out, int64_t batch, void *stream, const ExecFuse &f)
{
    if (!plan_first_pass_fusable(p)) return fail(MFFT_ERR_UNSUPPORTED, "first pass takes no fused input");
    if (batch < 1 || batch > p->batch) return fail(MFFT_ERR_INVALID, "batch %lld outside 1..%lld", (long long)batch, p->batch);
    return exec_all(p, in, out, stream, nullptr, batch, nullptr, &f);
}

}  // namespace mfft
