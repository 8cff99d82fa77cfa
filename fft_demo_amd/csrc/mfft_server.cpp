// Continuous-batching request server for mixed 1D FFT / 2D FFT / conv2d requests
// (C ABI in include/mfft.h). The reference has no serving layer; its nearest idea is the
// multi-stream host-staged batch split of fft_2d_c2c (tensorcore_fft_conv2d.cpp:79-146).
//
// One scheduler thread:
//   - takes newly submitted requests (submit never blocks on the GPU);
//   - retires finished work (hipEventQuery), freeing its HBM footprint;
//   - admits pending requests in arrival order while their footprint fits the HBM budget
//     (requests that do not fit are passed over, smaller ones behind them go first);
//   - coalesces the admitted small 1D transforms (n <= 4096) into ONE launch of the
//     work-queue kernel, and sends every other request to the least-loaded stream, so
//     requests that arrive while others are on the GPU overlap with them on the device.
// Plans, twiddle tables and filter spectra are built once and stay resident; each stream
// owns a workspace slab, so plans are shared by all streams (mfft_exec_ex).
// Plain mode (the baseline): fp32 plans, one request at a time on one stream, each
// completed before the next starts.
#include "../../include/mfft.h"

#include <hip/hip_runtime.h>
#include <sys/prctl.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

namespace {

thread_local std::string s_err;

int sfail(int code, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    s_err = buf;
    return code;
}

double now_us()
{
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Req {
    mfft_request r{};
    double t_submit = 0, t_start = 0;
    long long need = 0;
    int stream = -1;
    long long work = 0;       // counted in its stream's queued work while in flight
    hipEvent_t ev = nullptr;  // shared by the requests of one coalesced launch
    bool owns_ev = false;
    double out_scale = 1.0;
    int status = 0;
    bool admitted = false;
};

struct FilterHost {
    int id, fh, fw;
    std::vector<float> w;
};

}  // namespace

struct mfft_server_s {
    mfft_server_desc d{};
    int prec = MFFT_PREC_SPLIT_FP16;
    int fmt = MFFT_FMT_C16;
    std::vector<hipStream_t> streams;
    std::vector<void *> slabs;
    std::vector<int> outstanding;  // requests in flight per stream
    std::vector<long long> queued_work;  // complex points of the requests in flight per stream
    std::vector<hipEvent_t> free_events;  // completion events for reuse
    long long slab_bytes = 0;
    std::atomic<long long> static_bytes{0};  // slabs + conv operators + spectra
    mfft_batcher batcher[2] = {nullptr, nullptr};  // forward, inverse
    std::map<std::tuple<int, int, int, int>, mfft_plan> plans;  // (kind, nx, ny, dir)
    std::vector<FilterHost> filters;
    std::vector<std::pair<int, int>> stale_classes;  // (fh & 1, fw & 1) changed since their operators were built; under mu
    struct mfft_server_convs *conv = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req *> incoming;
    std::deque<Req *> pending;
    std::vector<Req *> inflight;
    std::vector<mfft_completion> done;
    std::atomic<long long> used{0};  // request I/O of admitted, unfinished requests
    std::atomic<long long> peak{0};  // largest static_bytes + used seen at an admission
    long long submitted = 0, completed = 0;  // under mu
    // written by the scheduler thread, read by mfft_server_get_stats
    std::atomic<long long> launches{0}, coalesced{0}, paired{0};
    std::atomic<long long> busy_ns{0};  // scheduler time spent retiring, admitting, launching
    std::atomic<bool> stop{false};
    int device = 0;
    std::thread sched;
};

namespace {

// Transform lengths the library plans: n >= 2 with prime factors in {2, 3, 5, 17}.
bool smooth_length(int n)
{
    if (n < 2) return false;
    for (int p : {2, 3, 5, 17})
        while (n % p == 0) n /= p;
    return n == 1;
}

long long io_bytes(const mfft_server_s *S, const mfft_request &r)
{
    const int eb = S->fmt == MFFT_FMT_C32 ? 8 : 4;
    if (r.kind == MFFT_REQ_FFT1D) return 2LL * r.n * r.batch * eb;
    if (r.kind == MFFT_REQ_FFT2D) return 2LL * r.nx * r.ny * r.batch * eb;
    return 2LL * r.nx * r.ny * 4;  // conv: fp32 image in and out
}

mfft_plan get_plan(mfft_server_s *S, int kind, int nx, int ny, int dir)
{
    auto key = std::make_tuple(kind, nx, ny, dir);
    auto it = S->plans.find(key);
    if (it != S->plans.end()) return it->second;
    mfft_desc d;
    mfft_desc_default(&d);
    d.precision = S->prec;
    d.direction = dir;
    d.in_fmt = d.out_fmt = S->fmt;
    d.scale = S->prec == MFFT_PREC_FP32 ? MFFT_SCALE_NONE : MFFT_SCALE_STAGE;
    d.workspace_limit = S->slab_bytes;
    d.external_workspace = 1;  // the streams' slabs
    if (!S->d.plain) d.mid_fmt = S->d.mid_fmt;
    mfft_plan p = nullptr;
    int rc = kind == MFFT_REQ_FFT1D ? mfft_plan_create_1d(&p, nx, S->d.max_batch_1d, &d)
                                    : mfft_plan_create_2d(&p, nx, ny, S->d.max_batch_2d, &d);
    if (rc) return nullptr;
    mfft_info inf;
    if (mfft_plan_info(p, &inf) || inf.workspace_bytes > S->slab_bytes) {
        // one transform's intermediates do not fit a stream's slab
        mfft_plan_destroy(p);
        return nullptr;
    }
    S->plans[key] = p;
    return p;
}

// Streams form two lanes: requests below kBulkPoints complex points go to the fast lane,
// larger ones to the bulk lane, so a small request never queues behind a large one on
// its stream. Within a lane, the stream with the least work in flight (complex points of
// its requests; a conv request counts its image twice, for the forward and the inverse
// transform), so a request does not wait behind a large one while the lane's other
// stream has little queued.
constexpr long long kBulkPoints = 1LL << 20;

long long req_points(const mfft_request &r)
{
    if (r.kind == MFFT_REQ_FFT1D) return (long long)r.n * r.batch;
    return (long long)r.nx * r.ny * (r.kind == MFFT_REQ_FFT2D ? r.batch : 1);
}

long long req_work(const mfft_request &r)
{
    return req_points(r) * (r.kind == MFFT_REQ_CONV2D ? 2 : 1);
}

// Filters whose sizes have the same parity share a frame (the class's largest height and
// width): a filter placed at the centre of a larger zero frame gives the same 'same'-mode
// result, so ONE conv operator per image shape and parity class holds every filter spectrum
// of the class. Each operator is bound to one stream (its workspace is its own).
struct ConvSlot {
    mfft_conv2d op = nullptr;
    std::vector<int> ids;
    int stream = 0;
};

const FilterHost *find_filter(const mfft_server_s *S, int id);

}  // namespace

struct mfft_server_convs {
    std::map<std::tuple<int, int, int, int>, ConvSlot> slots;  // (h, w, fh & 1, fw & 1)
    int next_stream = 0;
};

namespace {

ConvSlot *get_conv(mfft_server_s *S, int h, int w, int filter_id, int *index)
{
    const FilterHost *f = find_filter(S, filter_id);
    if (!f) return nullptr;
    const int ph = f->fh & 1, pw = f->fw & 1;
    auto key = std::make_tuple(h, w, ph, pw);
    auto it = S->conv->slots.find(key);
    if (it == S->conv->slots.end()) {
        int Fh = 0, Fw = 0;
        std::vector<const FilterHost *> cls;
        for (auto &g : S->filters)
            if ((g.fh & 1) == ph && (g.fw & 1) == pw) {
                cls.push_back(&g);
                Fh = std::max(Fh, g.fh);
                Fw = std::max(Fw, g.fw);
            }
        ConvSlot slot;
        std::vector<float> stack(cls.size() * (size_t)Fh * Fw, 0.f);
        for (size_t k = 0; k < cls.size(); ++k) {
            const FilterHost *g = cls[k];
            const int oy = (Fh - g->fh) / 2, ox = (Fw - g->fw) / 2;
            for (int y = 0; y < g->fh; ++y)
                for (int x = 0; x < g->fw; ++x)
                    stack[(k * Fh + oy + y) * (size_t)Fw + ox + x] = g->w[(size_t)y * g->fw + x];
            slot.ids.push_back(g->id);
        }
        mfft_conv2d_desc cd;
        mfft_conv2d_desc_default(&cd);
        cd.batch = 1;
        cd.channels = 1;
        cd.filters = (int)cls.size();
        cd.h = h;
        cd.w = w;
        cd.fh = Fh;
        cd.fw = Fw;
        cd.mode = MFFT_CONV_SAME;
        cd.precision = S->prec;
        // fp16 intermediates asked for the FFT plans: the conv operators keep fp16 spectra too
        cd.spec_fp16 = !S->d.plain && S->d.mid_fmt == MFFT_FMT_C16 ? 1 : 0;
        cd.workspace_limit = S->slab_bytes;
        {
            mfft_request probe{};
            probe.kind = MFFT_REQ_CONV2D;
            probe.nx = h;
            probe.ny = w;
            const bool bulk = req_points(probe) >= kBulkPoints;
            const int n = (int)S->streams.size(), nf = n > 1 ? n / 2 : n;
            const int lo = bulk && n > 1 ? nf : 0, width = bulk && n > 1 ? n - nf : nf;
            slot.stream = lo + S->conv->next_stream++ % width;
        }
        if (mfft_conv2d_create(&slot.op, &cd)) return nullptr;
        void *dfilt = nullptr;
        if (hipMalloc(&dfilt, stack.size() * sizeof(float)) != hipSuccess) {
            mfft_conv2d_destroy(slot.op);
            return nullptr;
        }
        hipStream_t s = S->streams[slot.stream];
        bool ok = hipMemcpy(dfilt, stack.data(), stack.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
                  mfft_conv2d_set_filters(slot.op, dfilt, MFFT_REAL_F32, s) == MFFT_OK &&
                  hipStreamSynchronize(s) == hipSuccess;
        (void)hipFree(dfilt);
        if (!ok) {
            mfft_conv2d_destroy(slot.op);
            return nullptr;
        }
        mfft_conv2d_info inf;
        mfft_conv2d_info_get(slot.op, &inf);
        S->static_bytes += inf.device_bytes;
        it = S->conv->slots.emplace(key, slot).first;
    }
    const std::vector<int> &ids = it->second.ids;
    auto p = std::find(ids.begin(), ids.end(), filter_id);
    if (p == ids.end()) return nullptr;
    *index = (int)(p - ids.begin());
    return &it->second;
}

const FilterHost *find_filter(const mfft_server_s *S, int id)
{
    for (auto &f : S->filters)
        if (f.id == id) return &f;
    return nullptr;
}

int pick_stream(mfft_server_s *S, bool bulk)
{
    const int n = (int)S->streams.size();
    const int nf = n > 1 ? n / 2 : n;  // fast lane [0, nf), bulk lane [nf, n)
    int lo = 0, hi = nf;
    if (bulk && n > 1) {
        lo = nf;
        hi = n;
    }
    int best = lo;
    for (int i = lo + 1; i < hi; ++i)
        if (S->queued_work[i] < S->queued_work[best] ||
            (S->queued_work[i] == S->queued_work[best] && S->outstanding[i] < S->outstanding[best]))
            best = i;
    return best;
}

// q runs on stream st until it completes
void track(mfft_server_s *S, Req *q, int st)
{
    q->stream = st;
    q->work = req_work(q->r);
    S->outstanding[st]++;
    S->queued_work[st] += q->work;
}

hipEvent_t take_event(mfft_server_s *S)
{
    if (!S->free_events.empty()) {
        hipEvent_t e = S->free_events.back();
        S->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Launch one non-coalesced request on stream `st`.
int launch_one(mfft_server_s *S, Req *q, int st)
{
    hipStream_t s = S->streams[st];
    const mfft_request &r = q->r;
    int rc = MFFT_OK;
    if (r.kind == MFFT_REQ_FFT1D || r.kind == MFFT_REQ_FFT2D) {
        mfft_plan p = r.kind == MFFT_REQ_FFT1D ? get_plan(S, r.kind, r.n, 0, r.direction)
                                               : get_plan(S, r.kind, r.nx, r.ny, r.direction);
        if (!p) return sfail(MFFT_ERR_UNSUPPORTED, "no plan for this shape within the stream workspace (%lld bytes): %s", S->slab_bytes, mfft_last_error());
        mfft_info inf;
        mfft_plan_info(p, &inf);
        q->out_scale = inf.out_scale;
        rc = mfft_exec_ex(p, r.in, r.out, r.batch, S->slabs[st], s);
    } else {
        int idx = 0;
        ConvSlot *c = get_conv(S, r.nx, r.ny, r.filter_id, &idx);
        if (!c) return sfail(MFFT_ERR_UNSUPPORTED, "conv operator for filter %d", r.filter_id);
        st = c->stream;
        s = S->streams[st];
        rc = mfft_conv2d_forward_select(c->op, r.in, MFFT_REAL_F32, r.out, MFFT_REAL_F32, idx, s);
    }
    if (rc) return rc;
    q->ev = take_event(S);
    if (!q->ev) return MFFT_ERR_HIP;
    q->owns_ev = true;
    if (hipEventRecord(q->ev, s) != hipSuccess) return MFFT_ERR_HIP;
    track(S, q, st);
    S->launches++;
    return MFFT_OK;
}

// Two conv requests on one operator: one complex transform carries both images.
int launch_conv_two(mfft_server_s *S, Req *a, Req *b, ConvSlot *c, int ia, int ib)
{
    hipStream_t s = S->streams[c->stream];
    int rc = mfft_conv2d_forward_two(c->op, a->r.in, b->r.in, MFFT_REAL_F32, a->r.out, b->r.out, MFFT_REAL_F32, ia, ib, s);
    if (rc) return rc;
    hipEvent_t ev = take_event(S);
    if (!ev) return MFFT_ERR_HIP;
    if (hipEventRecord(ev, s) != hipSuccess) {
        S->free_events.push_back(ev);
        return MFFT_ERR_HIP;
    }
    a->ev = b->ev = ev;
    a->owns_ev = true;
    b->owns_ev = false;
    track(S, a, c->stream);
    track(S, b, c->stream);
    S->launches++;
    S->paired += 2;
    return MFFT_OK;
}

void complete(mfft_server_s *S, Req *q, int status)
{
    mfft_completion c;
    c.tag = q->r.tag;
    c.status = status;
    c.submit_us = q->t_submit;
    c.start_us = q->t_start;
    c.end_us = now_us();
    c.out_scale = q->out_scale;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->done.push_back(c);
        S->completed++;
    }
    if (q->admitted) S->used -= q->need;
    if (q->stream >= 0) {
        S->outstanding[q->stream]--;
        S->queued_work[q->stream] -= q->work;
    }
    if (q->owns_ev && q->ev) S->free_events.push_back(q->ev);
    delete q;
}

void scheduler(mfft_server_s *S)
{
    (void)hipSetDevice(S->device);
    // Completions are found by polling events every few microseconds while requests are
    // in flight; the default 50 us timer slack of a Linux thread would stretch each poll
    // period, and every request's latency with it, to about 55 us.
    (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0UL, 0UL, 0UL);
    const bool plain = S->d.plain != 0;
    std::vector<Req *> small;
    std::vector<mfft_batch_item> items;
    std::map<ConvSlot *, std::vector<std::pair<Req *, int>>> convq;  // admitted conv requests per operator
    std::vector<std::pair<int, int>> stale;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->incoming.empty() && S->pending.empty() && S->inflight.empty()) {
                if (S->stop) break;
                S->cv.wait_for(lk, std::chrono::microseconds(200));
            }
            while (!S->incoming.empty()) {
                S->pending.push_back(S->incoming.front());
                S->incoming.pop_front();
            }
            stale.swap(S->stale_classes);
        }
        // filters added since: drop the operators of their classes (nothing runs on them --
        // filters are only added while no request is outstanding) before anything is admitted
        for (auto &cls : stale)
            for (auto it = S->conv->slots.begin(); it != S->conv->slots.end();) {
                if (std::get<2>(it->first) == cls.first && std::get<3>(it->first) == cls.second) {
                    mfft_conv2d_info inf;
                    if (mfft_conv2d_info_get(it->second.op, &inf) == MFFT_OK) S->static_bytes -= inf.device_bytes;
                    mfft_conv2d_destroy(it->second.op);
                    it = S->conv->slots.erase(it);
                } else {
                    ++it;
                }
            }
        stale.clear();
        const auto t_busy = std::chrono::steady_clock::now();
        // retire: query each event once; requests of one coalesced or paired launch share
        // their event and retire together (its owner returns it to the pool)
        {
            std::vector<Req *> keep;
            std::vector<std::pair<Req *, hipError_t>> fin;
            std::map<hipEvent_t, hipError_t> seen;
            keep.reserve(S->inflight.size());
            for (Req *q : S->inflight) {
                auto f = seen.find(q->ev);
                hipError_t e = f != seen.end() ? f->second : (seen[q->ev] = hipEventQuery(q->ev));
                if (e == hipErrorNotReady)
                    keep.push_back(q);
                else
                    fin.push_back({q, e});
            }
            S->inflight.swap(keep);
            for (auto &d : fin) complete(S, d.first, d.second == hipSuccess ? MFFT_OK : MFFT_ERR_HIP);
        }
        // admit
        small.clear();
        items.clear();
        convq.clear();
        const long long budget = S->d.hbm_budget;
        for (auto it = S->pending.begin(); it != S->pending.end();) {
            if (plain && !S->inflight.empty()) break;
            Req *q = *it;
            q->need = io_bytes(S, q->r);
            if (S->static_bytes + S->used + q->need > budget) {
                if (q->need + S->static_bytes > budget) {  // can never fit: reject
                    it = S->pending.erase(it);
                    complete(S, q, MFFT_ERR_NOMEM);
                    continue;
                }
                ++it;
                continue;
            }
            q->t_start = now_us();
            q->admitted = true;
            S->used += q->need;
            S->peak = std::max(S->peak.load(), S->static_bytes + S->used);
            const bool is_small = !plain && q->r.kind == MFFT_REQ_FFT1D && q->r.n <= 4096 &&
                                  (int)small.size() < S->d.coalesce_max && q->r.direction == MFFT_FORWARD;
            if (is_small) {
                small.push_back(q);
                items.push_back(mfft_batch_item{q->r.in, q->r.out, q->r.n, q->r.batch});
                q->out_scale = 1.0 / q->r.n;
                it = S->pending.erase(it);
                continue;
            }
            if (!plain && q->r.kind == MFFT_REQ_CONV2D) {
                // held back to pair with another request on the same operator
                int idx = 0;
                ConvSlot *c = get_conv(S, q->r.nx, q->r.ny, q->r.filter_id, &idx);
                it = S->pending.erase(it);
                if (!c) {
                    complete(S, q, sfail(MFFT_ERR_UNSUPPORTED, "conv operator for filter %d", q->r.filter_id));
                    continue;
                }
                convq[c].push_back({q, idx});
                continue;
            }
            int st = plain ? 0 : pick_stream(S, req_points(q->r) >= kBulkPoints);
            int rc = launch_one(S, q, st);
            it = S->pending.erase(it);
            if (rc) {
                complete(S, q, rc);
                continue;
            }
            S->inflight.push_back(q);
            if (plain) break;
        }
        for (auto &kv : convq) {
            auto &v = kv.second;
            size_t k = 0;
            for (; k + 1 < v.size(); k += 2) {
                int rc = launch_conv_two(S, v[k].first, v[k + 1].first, kv.first, v[k].second, v[k + 1].second);
                if (rc) {
                    complete(S, v[k].first, rc);
                    complete(S, v[k + 1].first, rc);
                    continue;
                }
                S->inflight.push_back(v[k].first);
                S->inflight.push_back(v[k + 1].first);
            }
            if (k < v.size()) {
                Req *q = v[k].first;
                int rc = launch_one(S, q, kv.first->stream);
                if (rc)
                    complete(S, q, rc);
                else
                    S->inflight.push_back(q);
            }
        }
        if (!small.empty()) {
            int st = pick_stream(S, false);
            int rc = mfft_batcher_exec(S->batcher[0], items.data(), (int)items.size(), S->streams[st]);
            hipEvent_t ev = rc ? nullptr : take_event(S);
            if (ev && hipEventRecord(ev, S->streams[st]) == hipSuccess) {
                S->launches++;
                S->coalesced += (long long)small.size();
                for (size_t k = 0; k < small.size(); ++k) {
                    small[k]->ev = ev;
                    small[k]->owns_ev = k == 0;
                    track(S, small[k], st);
                    S->inflight.push_back(small[k]);
                }
            } else {
                if (ev) S->free_events.push_back(ev);
                for (Req *q : small) complete(S, q, rc ? rc : MFFT_ERR_HIP);
            }
        }
        S->busy_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t_busy)
                          .count();
        if (plain && !S->inflight.empty()) {
            (void)hipStreamSynchronize(S->streams[0]);  // one request at a time
        } else if (!S->inflight.empty() && S->pending.empty()) {
            // until the next poll of the in-flight events, or a submission (which notifies)
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->incoming.empty()) S->cv.wait_for(lk, std::chrono::microseconds(5));
        }
    }
}

}  // namespace

extern "C" {

int mfft_server_desc_default(mfft_server_desc *d)
{
    if (!d) return MFFT_ERR_INVALID;
    memset(d, 0, sizeof *d);
    d->hbm_budget = 64LL << 30;
    d->streams = 4;
    d->precision = MFFT_PREC_SPLIT_FP16;
    d->max_batch_1d = 64;
    d->max_batch_2d = 16;
    d->coalesce_max = 256;
    d->slab_bytes = 512LL << 20;
    d->mid_fmt = -1;
    return MFFT_OK;
}

int mfft_server_create(mfft_server *out, const mfft_server_desc *d)
{
    if (!out || !d) return sfail(MFFT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (d->streams < 1 || d->streams > 16 || d->hbm_budget <= 0 || d->max_batch_1d < 1 || d->max_batch_2d < 1)
        return sfail(MFFT_ERR_INVALID, "bad descriptor");
    mfft_server S = new mfft_server_s;
    S->conv = new mfft_server_convs;
    S->d = *d;
    (void)hipGetDevice(&S->device);
    if (S->d.plain) {
        S->d.streams = 1;
        S->prec = MFFT_PREC_FP32;
        S->fmt = MFFT_FMT_C32;
    } else {
        S->prec = d->precision;
        S->fmt = d->precision == MFFT_PREC_FP32 ? MFFT_FMT_C32 : d->precision == MFFT_PREC_BF16 ? MFFT_FMT_CBF16 : MFFT_FMT_C16;
    }
    S->slab_bytes = d->slab_bytes > 0 ? d->slab_bytes : (512LL << 20);
    S->streams.resize(S->d.streams);
    S->slabs.resize(S->d.streams, nullptr);
    S->outstanding.assign(S->d.streams, 0);
    S->queued_work.assign(S->d.streams, 0);
    for (int i = 0; i < S->d.streams; ++i) {
        if (hipStreamCreateWithFlags(&S->streams[i], hipStreamNonBlocking) != hipSuccess ||
            hipMalloc(&S->slabs[i], S->slab_bytes) != hipSuccess) {
            mfft_server_destroy(S);
            return sfail(MFFT_ERR_NOMEM, "streams / workspace slabs");
        }
        S->static_bytes += S->slab_bytes;
    }
    if (!S->d.plain) {
        mfft_desc bd;
        mfft_desc_default(&bd);
        bd.precision = S->prec;
        bd.in_fmt = bd.out_fmt = S->fmt;
        bd.scale = MFFT_SCALE_STAGE;
        if (mfft_batcher_create(&S->batcher[0], &bd, S->d.coalesce_max)) {
            std::string m = mfft_last_error();
            mfft_server_destroy(S);
            return sfail(MFFT_ERR_HIP, "%s", m.c_str());
        }
    }
    S->sched = std::thread(scheduler, S);
    *out = S;
    return MFFT_OK;
}

int mfft_server_add_filter(mfft_server S, int filter_id, const float *weights, int fh, int fw)
{
    if (!S || !weights || fh < 1 || fw < 1) return sfail(MFFT_ERR_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->incoming.empty() || !S->pending.empty() || S->submitted != S->completed)
        return sfail(MFFT_ERR_INVALID, "filters are registered while no request is outstanding");
    FilterHost f{filter_id, fh, fw, std::vector<float>(weights, weights + (size_t)fh * fw)};
    // A conv operator already built for this filter's size-parity class holds the spectra
    // of the class as it was: the scheduler drops those operators before its next
    // admission, and the next request of the class rebuilds them with every filter.
    S->stale_classes.push_back({fh & 1, fw & 1});
    auto old = std::find_if(S->filters.begin(), S->filters.end(), [&](const FilterHost &g) { return g.id == filter_id; });
    if (old != S->filters.end()) {
        S->stale_classes.push_back({old->fh & 1, old->fw & 1});  // a re-registered id leaves its old class
        *old = f;
    } else {
        S->filters.push_back(f);
    }
    return MFFT_OK;
}

int mfft_server_submit(mfft_server S, const mfft_request *r)
{
    if (!S || !r || !r->in || !r->out) return sfail(MFFT_ERR_INVALID, "null argument");
    if (r->kind < MFFT_REQ_FFT1D || r->kind > MFFT_REQ_CONV2D) return sfail(MFFT_ERR_INVALID, "kind");
    if ((r->kind == MFFT_REQ_FFT1D && (r->batch < 1 || r->batch > S->d.max_batch_1d)) ||
        (r->kind == MFFT_REQ_FFT2D && (r->batch < 1 || r->batch > S->d.max_batch_2d)))
        return sfail(MFFT_ERR_INVALID, "batch outside the server's plans");
    if (r->direction != 0 && r->direction != MFFT_FORWARD && r->direction != MFFT_INVERSE)
        return sfail(MFFT_ERR_INVALID, "direction %d", r->direction);
    // Shapes are checked here, before the request can join a coalesced launch: the batcher
    // rejects a whole launch at its first unsupported length, which would fail every small
    // request admitted with it.
    if (r->kind == MFFT_REQ_FFT1D && !smooth_length(r->n))
        return sfail(MFFT_ERR_INVALID, "1D length %d: not 2^a 3^b 5^c 17^d >= 2", r->n);
    if (r->kind == MFFT_REQ_FFT2D && (!smooth_length(r->nx) || !smooth_length(r->ny)))
        return sfail(MFFT_ERR_INVALID, "2D shape %d x %d: sides must be 2^a 3^b 5^c 17^d >= 2", r->nx, r->ny);
    if (r->kind == MFFT_REQ_CONV2D) {
        if (r->nx < 1 || r->ny < 1) return sfail(MFFT_ERR_INVALID, "conv image %d x %d", r->nx, r->ny);
        std::lock_guard<std::mutex> lk(S->mu);
        if (!find_filter(S, r->filter_id)) return sfail(MFFT_ERR_INVALID, "filter %d is not registered", r->filter_id);
    }
    Req *q = new Req;
    q->r = *r;
    if (q->r.direction == 0) q->r.direction = MFFT_FORWARD;
    q->t_submit = now_us();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->incoming.push_back(q);
        S->submitted++;
    }
    S->cv.notify_one();
    return MFFT_OK;
}

int mfft_server_poll(mfft_server S, mfft_completion *out, int max, int *n)
{
    if (!S || !n || (max > 0 && !out)) return sfail(MFFT_ERR_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    int k = std::min<int>(max, (int)S->done.size());
    for (int i = 0; i < k; ++i) out[i] = S->done[i];
    S->done.erase(S->done.begin(), S->done.begin() + k);
    *n = k;
    return MFFT_OK;
}

int mfft_server_wait_idle(mfft_server S, double timeout_s)
{
    if (!S) return sfail(MFFT_ERR_INVALID, "null server");
    double t0 = now_us();
    while (true) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->submitted == S->completed) return MFFT_OK;
        }
        if (timeout_s > 0 && now_us() - t0 > timeout_s * 1e6) return sfail(MFFT_ERR_INVALID, "timeout");
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
}

int mfft_server_get_stats(mfft_server S, mfft_server_stats *st)
{
    if (!S || !st) return sfail(MFFT_ERR_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    st->submitted = S->submitted;
    st->completed = S->completed;
    st->launches = S->launches;
    st->coalesced = S->coalesced;
    st->paired = S->paired;
    st->static_bytes = S->static_bytes;
    st->inflight_bytes = S->used;
    st->peak_bytes = S->peak;
    st->busy_us = S->busy_ns / 1000;
    return MFFT_OK;
}

int mfft_server_destroy(mfft_server S)
{
    if (!S) return MFFT_OK;
    S->stop = true;
    S->cv.notify_all();
    if (S->sched.joinable()) S->sched.join();
    (void)hipDeviceSynchronize();
    if (S->conv)
        for (auto &kv : S->conv->slots) mfft_conv2d_destroy(kv.second.op);
    delete S->conv;
    for (auto &kv : S->plans) mfft_plan_destroy(kv.second);
    for (auto &b : S->batcher)
        if (b) mfft_batcher_destroy(b);
    for (size_t i = 0; i < S->streams.size(); ++i) {
        if (S->slabs[i]) (void)hipFree(S->slabs[i]);
        if (S->streams[i]) (void)hipStreamDestroy(S->streams[i]);
    }
    for (hipEvent_t e : S->free_events) (void)hipEventDestroy(e);
    for (Req *q : S->pending) delete q;
    for (Req *q : S->incoming) delete q;
    delete S;
    return MFFT_OK;
}

const char *mfft_server_last_error(void) { return s_err.c_str(); }

}  // extern "C"
