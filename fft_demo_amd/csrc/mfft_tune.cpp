// Schedule autotuner (C ABI in include/mfft.h). The reference fixes its schedule per size in
// a switch (tcfftCreate, tcfft_half.cu:1250-1330; tcfftCreate_2d, tcfft_half_2d.cu:1167-1206);
// here the planner proposes, and this tuner times the candidates on the device and keeps the
// fastest per (device arch, shape, batch, precision, formats, direction, input order).
//
// What is tuned is the schedule only -- never the precision: the first-pass length of a
// four-step split (split_hint; 1 = one mixed-radix pass instead of the odd-factor split) and
// the batch chunk (workspace_limit: how many transforms' intermediates are in flight between
// the passes, i.e. how much of the intermediate can stay in the 256 MB last-level cache).
// The intermediate format (mfft_desc.mid_fmt) changes rounding and stays the caller's choice.
//
// Results live in a process-wide cache, saved to and loaded from a text file, and a plan
// created with split_hint = MFFT_SPLIT_TUNED takes its split and chunk from the cache.
#include "../../include/mfft.h"
#include "mfft_fast.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace mfft {
hipError_t gen_twiddle(float2 *out, long long n, long long count, long long mult, int sign, hipStream_t s,
                       double scale = 1.0);
}

namespace {

thread_local std::string t_err;

int tfail(int code, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}

using Key = std::tuple<std::string, int, int, long long, int, int, int, int, int, int>;

struct Entry {
    int split_hint = 0;
    long long workspace_limit = 0;
    double ms = 0, default_ms = 0;
    int candidates = 0;
};

std::mutex g_mu;
std::map<Key, Entry> g_cache;

std::string current_arch()
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "none";
    std::string a = prop.gcnArchName;
    size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

Key make_key(const std::string &arch, int nx, int ny, long long batch, const mfft_desc *d)
{
    return Key(arch, nx, ny, batch, d->precision, d->in_fmt, d->out_fmt, d->mid_fmt, d->direction, d->in_order);
}

int fmt_bytes(int f) { return f == MFFT_FMT_C32 ? 8 : 4; }

bool is_pow2(long long n) { return n > 0 && (n & (n - 1)) == 0; }

long long odd_part(long long n)
{
    while (n > 0 && n % 2 == 0) n /= 2;
    return n;
}

// Time `reps` execs of a plan; returns the mean ms per exec (or a negative value on error).
double time_plan(mfft_plan p, void *in, void *out, hipStream_t s, int reps)
{
    for (int i = 0; i < 2; ++i)
        if (mfft_exec(p, in, out, s)) return -1;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1;
    }
    double ms = -1;
    if (hipEventRecord(a, s) == hipSuccess) {
        bool ok = true;
        for (int i = 0; i < reps && ok; ++i) ok = mfft_exec(p, in, out, s) == MFFT_OK;
        float t = 0;
        if (ok && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
            hipEventElapsedTime(&t, a, b) == hipSuccess)
            ms = t / reps;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

}  // namespace

extern "C" {

int mfft_tune_lookup(const char *arch, int nx, int ny, int64_t batch, const mfft_desc *d, mfft_tune_entry *out)
{
    if (!d || !out) return tfail(MFFT_ERR_INVALID, "null argument");
    std::string a = arch ? std::string(arch) : current_arch();
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(make_key(a, nx, ny, batch, d));
    if (it == g_cache.end()) return tfail(MFFT_ERR_UNSUPPORTED, "no tuned schedule for this plan");
    out->split_hint = it->second.split_hint;
    out->workspace_limit = it->second.workspace_limit;
    out->ms = it->second.ms;
    out->default_ms = it->second.default_ms;
    out->candidates = it->second.candidates;
    return MFFT_OK;
}

int mfft_autotune(int nx, int ny, int64_t batch, const mfft_desc *d, int reps, mfft_tune_entry *out)
{
    if (!d || !out || nx < 2 || ny < 0 || ny == 1 || batch < 1) return tfail(MFFT_ERR_INVALID, "bad argument");
    if (d->split_hint == MFFT_SPLIT_TUNED) return tfail(MFFT_ERR_INVALID, "tune from an explicit descriptor");
    if (reps < 1) reps = 5;
    const bool two_d = ny > 0;
    const long long n = two_d ? (long long)nx * ny : nx;
    // candidate split points (1D): the planner's choice, every power-of-two first pass of a
    // power-of-two length (the rest in one pass when it fits a tile, else split again), and
    // for odd * 2^a lengths the single mixed-radix pass besides the odd-factor split
    std::vector<int> splits{0};
    if (!two_d && d->in_order == MFFT_ORDER_NATURAL) {
        if (is_pow2(n) && n > 4096)
            for (int l1 = 2; l1 <= mfft::kColsMaxR; l1 *= 2)
                if (n / l1 >= 16 && n % l1 == 0) splits.push_back(l1);
        if (odd_part(n) > 1) splits.push_back(1);
    }
    if (two_d && odd_part(nx) * odd_part(ny) > 1 && nx <= 4096 && ny <= 4096) splits.push_back(1);
    // 2D: the first-pass length of a split column axis
    if (two_d && d->in_order == MFFT_ORDER_NATURAL && is_pow2(nx) && nx > 256)
        for (int l1 = 2; l1 <= mfft::kColsMaxR; l1 *= 2)
            if (nx / l1 >= (l1 > 256 ? 2 : 16) && nx % l1 == 0) splits.push_back(l1);
    // candidate chunks: 16 MiB .. 1 GiB of intermediates in flight between the passes (at
    // least one transform, at most the whole batch), capped by the caller's limit
    const long long per = n * 8;
    std::vector<long long> limits{0};
    for (long long m = 16LL << 20; m <= (1LL << 30); m *= 2) {
        if (d->workspace_limit > 0 && m > d->workspace_limit) break;
        if (m < per) continue;
        limits.push_back(m);
        if (m >= per * batch) break;
    }
    hipStream_t s = nullptr;
    void *in = nullptr, *outb = nullptr;
    const size_t ib = (size_t)n * batch * fmt_bytes(d->in_fmt), ob = (size_t)n * batch * fmt_bytes(d->out_fmt);
    if (hipMalloc(&in, ib) != hipSuccess || hipMalloc(&outb, ob) != hipSuccess || hipStreamCreate(&s) != hipSuccess) {
        (void)hipFree(in);
        (void)hipFree(outb);
        return tfail(MFFT_ERR_NOMEM, "tuning buffers (%zu + %zu bytes)", ib, ob);
    }
    // representative input (all-zero data runs measurably faster): points of the unit circle
    // W_1009^(37 i), generated on the device and converted to the plan's input format
    {
        bool ok = true;
        if (d->in_fmt == MFFT_FMT_C32) {
            ok = mfft::gen_twiddle((float2 *)in, 1009, n * batch, 37, -1, s) == hipSuccess;
        } else {
            void *tmp = nullptr;
            ok = hipMalloc(&tmp, (size_t)n * batch * 8) == hipSuccess &&
                 mfft::gen_twiddle((float2 *)tmp, 1009, n * batch, 37, -1, s) == hipSuccess &&
                 mfft_stage_convert(tmp, MFFT_FMT_C32, in, d->in_fmt, n * batch, 1, nullptr, 0, nullptr, 0, 1.0, s) ==
                     MFFT_OK &&
                 hipStreamSynchronize(s) == hipSuccess;
            (void)hipFree(tmp);
        }
        if (!ok || hipStreamSynchronize(s) != hipSuccess) {
            (void)hipStreamDestroy(s);
            (void)hipFree(in);
            (void)hipFree(outb);
            return tfail(MFFT_ERR_HIP, "tuning input");
        }
    }
    Entry best;
    best.ms = 1e30;
    double def_ms = -1;
    int tried = 0;
    for (int sp : splits)
        for (long long lim : limits) {
            mfft_desc c = *d;
            c.split_hint = sp;
            c.workspace_limit = lim;
            c.external_workspace = 0;
            mfft_plan p = nullptr;
            int rc = two_d ? mfft_plan_create_2d(&p, nx, ny, batch, &c) : mfft_plan_create_1d(&p, nx, batch, &c);
            if (rc) continue;  // the planner does not take this candidate
            double ms = time_plan(p, in, outb, s, reps);
            mfft_plan_destroy(p);
            if (ms < 0) continue;
            ++tried;
            if (sp == 0 && lim == 0) def_ms = ms;
            if (ms < best.ms) {
                best.ms = ms;
                best.split_hint = sp;
                best.workspace_limit = lim;
            }
        }
    (void)hipStreamDestroy(s);
    (void)hipFree(in);
    (void)hipFree(outb);
    if (!tried) return tfail(MFFT_ERR_UNSUPPORTED, "no candidate schedule ran");
    best.default_ms = def_ms;
    best.candidates = tried;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cache[make_key(current_arch(), nx, two_d ? ny : 0, batch, d)] = best;
    }
    out->split_hint = best.split_hint;
    out->workspace_limit = best.workspace_limit;
    out->ms = best.ms;
    out->default_ms = best.default_ms;
    out->candidates = best.candidates;
    return MFFT_OK;
}

int mfft_tune_cache_save(const char *path)
{
    if (!path) return tfail(MFFT_ERR_INVALID, "null path");
    FILE *f = fopen(path, "w");
    if (!f) return tfail(MFFT_ERR_INVALID, "cannot write %s", path);
    std::lock_guard<std::mutex> lk(g_mu);
    fprintf(f, "mfft-tune 1\n");
    for (auto &kv : g_cache) {
        const Key &k = kv.first;
        const Entry &e = kv.second;
        fprintf(f, "%s %d %d %lld %d %d %d %d %d %d : %d %lld %.6f %.6f %d\n", std::get<0>(k).c_str(), std::get<1>(k),
                std::get<2>(k), std::get<3>(k), std::get<4>(k), std::get<5>(k), std::get<6>(k), std::get<7>(k),
                std::get<8>(k), std::get<9>(k), e.split_hint, e.workspace_limit, e.ms, e.default_ms, e.candidates);
    }
    fclose(f);
    return MFFT_OK;
}

int mfft_tune_cache_load(const char *path)
{
    if (!path) return tfail(MFFT_ERR_INVALID, "null path");
    FILE *f = fopen(path, "r");
    if (!f) return tfail(MFFT_ERR_INVALID, "cannot read %s", path);
    char line[512];
    int version = 0;
    if (!fgets(line, sizeof line, f) || sscanf(line, "mfft-tune %d", &version) != 1 || version != 1) {
        fclose(f);
        return tfail(MFFT_ERR_INVALID, "%s: not a tune cache (version 1)", path);
    }
    std::map<Key, Entry> add;
    int lineno = 1;
    while (fgets(line, sizeof line, f)) {
        ++lineno;
        char arch[64];
        int nx, ny, prec, inf, outf, mid, dir, order;
        long long batch;
        Entry e;
        if (sscanf(line, "%63s %d %d %lld %d %d %d %d %d %d : %d %lld %lf %lf %d", arch, &nx, &ny, &batch, &prec, &inf,
                   &outf, &mid, &dir, &order, &e.split_hint, &e.workspace_limit, &e.ms, &e.default_ms,
                   &e.candidates) != 15) {
            fclose(f);
            return tfail(MFFT_ERR_INVALID, "%s:%d: malformed entry", path, lineno);
        }
        add[Key(arch, nx, ny, batch, prec, inf, outf, mid, dir, order)] = e;
    }
    fclose(f);
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &kv : add) g_cache[kv.first] = kv.second;
    return MFFT_OK;
}

int mfft_tune_cache_clear(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache.clear();
    return MFFT_OK;
}

const char *mfft_tune_last_error(void) { return t_err.c_str(); }

}  // extern "C"
