// The reference's plan/exec and harness API (tcfft_half.h, tcfft_half_2d.h, cufft.h) over
// the MI355X library. See include/tcfft_half.h for the mapping of every entry point.
#include "../../include/mfft.h"
#include "../../include/tcfft_half.h"
#include "../../include/tcfft_half_2d.h"
#include "mfft_internal.h"

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local int g_compat_err = 0;

int set_err(int code)
{
    g_compat_err = code;
    return code;
}

// radix list of the digit-reversed order: the reference's list for every size tcfftCreate
// knows (tcfft_half.cu:1250-1330), the planner's rule for the rest
std::vector<int> radix_list(int n)
{
    std::vector<int> r;
    mfft::digit_reversal_radices(n, r);
    return r;
}

int env_precision()
{
    const char *e = getenv("MFFT_TCFFT_PRECISION");
    if (e && std::string(e) == "split") return MFFT_PREC_SPLIT_FP16;
    return MFFT_PREC_FP16;
}

// A reference-style plan: digit-reversed input fused into the first pass when the
// planner can, else a device gather back to natural order before the natural plan.
struct Compat {
    mfft_plan plan = nullptr;
    int nx = 0, ny = 0;  // 1D: nx = n, ny = 0
    long long batch = 0;
    bool fused = true;
    int *inv_x = nullptr, *inv_y = nullptr;  // inverse permutations (device)
    void *tmp = nullptr;
    ~Compat()
    {
        if (plan) mfft_plan_destroy(plan);
        (void)hipFree(inv_x);
        (void)hipFree(inv_y);
        (void)hipFree(tmp);
    }
};

mfft_desc compat_desc(int prec, int order)
{
    mfft_desc d;
    mfft_desc_default(&d);
    d.precision = prec;
    d.scale = MFFT_SCALE_NONE;  // the reference's unnormalised contract
    d.in_fmt = d.out_fmt = prec == MFFT_PREC_FP32 ? MFFT_FMT_C32 : MFFT_FMT_C16;
    d.in_order = order;
    return d;
}

int upload_inverse(const std::vector<int> &rev, int **dev)
{
    std::vector<int> inv(rev.size());
    for (size_t i = 0; i < rev.size(); ++i) inv[rev[i]] = (int)i;
    if (hipMalloc(dev, sizeof(int) * inv.size()) != hipSuccess) return MFFT_ERR_NOMEM;
    if (hipMemcpy(*dev, inv.data(), sizeof(int) * inv.size(), hipMemcpyHostToDevice) != hipSuccess)
        return MFFT_ERR_HIP;
    return MFFT_OK;
}

Compat *make_compat(int nx, int ny, long long batch, int prec, bool rev_order)
{
    Compat *c = new Compat;
    c->nx = nx;
    c->ny = ny;
    c->batch = batch;
    mfft_desc d = compat_desc(prec, rev_order ? MFFT_ORDER_DIGIT_REVERSED : MFFT_ORDER_NATURAL);
    int rc = ny ? mfft_plan_create_2d(&c->plan, nx, ny, batch, &d) : mfft_plan_create_1d(&c->plan, nx, batch, &d);
    if (rc == MFFT_ERR_UNSUPPORTED && rev_order) {
        // the planner fuses the permutation only into single-pass axes: gather instead
        d.in_order = MFFT_ORDER_NATURAL;
        rc = ny ? mfft_plan_create_2d(&c->plan, nx, ny, batch, &d) : mfft_plan_create_1d(&c->plan, nx, batch, &d);
        if (rc == MFFT_OK) {
            c->fused = false;
            std::vector<int> rx(nx), ry(ny > 0 ? ny : 1);
            rc = mfft_input_permutation(c->plan, 0, rx.data(), nx);
            if (!rc && ny) rc = mfft_input_permutation(c->plan, 1, ry.data(), ny);
            if (!rc) rc = upload_inverse(rx, &c->inv_x);
            if (!rc && ny) rc = upload_inverse(ry, &c->inv_y);
            size_t bytes = (size_t)batch * nx * (ny ? ny : 1) * (prec == MFFT_PREC_FP32 ? 8 : 4);
            if (!rc && hipMalloc(&c->tmp, bytes) != hipSuccess) rc = MFFT_ERR_NOMEM;
        }
    }
    if (rc) {
        set_err(rc);
        delete c;
        return nullptr;
    }
    set_err(MFFT_OK);
    return c;
}

int compat_exec(Compat *c, void *data, hipStream_t s)
{
    if (!c) return set_err(MFFT_ERR_INVALID);
    const int fmt = MFFT_FMT_C16;
    if (c->fused) return set_err(mfft_exec(c->plan, data, data, s));
    long long inner = (long long)c->nx * (c->ny ? c->ny : 1);
    // natural[b][x][y] = data[b][inv_x[x]][inv_y[y]]
    int rc = mfft_stage_convert(data, fmt, c->tmp, fmt, inner, c->batch, c->inv_x, c->nx, c->inv_y, c->ny, 1.0, s);
    if (!rc) rc = mfft_exec(c->plan, c->tmp, data, s);
    return set_err(rc);
}

// ---------------------------------------------------------------- harness state
struct Harness {
    Compat *c = nullptr;
    void *d_data = nullptr;   // device transform buffer
    void *d_f64 = nullptr;    // device fp64 staging buffer
    int nx = 0, ny = 0;
    long long batch = 0;
    int fmt = MFFT_FMT_C16;
    void reset()
    {
        delete c;
        c = nullptr;
        (void)hipFree(d_data);
        (void)hipFree(d_f64);
        d_data = d_f64 = nullptr;
    }
};

Harness g_h1, g_h2, g_hcu, g_hcu2;

void harness_setup(Harness &h, const double *data, int nx, int ny, long long batch, int prec)
{
    h.reset();
    h.nx = nx;
    h.ny = ny;
    h.batch = batch;
    h.fmt = prec == MFFT_PREC_FP32 ? MFFT_FMT_C32 : MFFT_FMT_C16;
    h.c = make_compat(nx, ny, batch, prec, false);
    if (!h.c) return;
    long long inner = (long long)nx * (ny ? ny : 1), total = inner * batch;
    if (hipMalloc(&h.d_data, total * (h.fmt == MFFT_FMT_C32 ? 8 : 4)) != hipSuccess ||
        hipMalloc(&h.d_f64, total * 16) != hipSuccess) {
        set_err(MFFT_ERR_NOMEM);
        return;
    }
    if (hipMemcpy(h.d_f64, data, total * 16, hipMemcpyHostToDevice) != hipSuccess) {
        set_err(MFFT_ERR_HIP);
        return;
    }
    int rc = mfft_stage_convert(h.d_f64, MFFT_FMT_C64, h.d_data, h.fmt, inner, batch, nullptr, 0, nullptr, 0, 1.0, 0);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = MFFT_ERR_HIP;
    set_err(rc);
}

void harness_doit(Harness &h, int iter)
{
    if (!h.c) {
        set_err(MFFT_ERR_INVALID);
        return;
    }
    int rc = MFFT_OK;
    for (int t = 0; t < iter && !rc; ++t) rc = mfft_exec(h.c->plan, h.d_data, h.d_data, 0);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = MFFT_ERR_HIP;
    set_err(rc);
}

void harness_finalize(Harness &h, double *result, long long nbatch)
{
    if (!h.c) {
        set_err(MFFT_ERR_INVALID);
        return;
    }
    long long inner = (long long)h.nx * (h.ny ? h.ny : 1);
    long long nb = nbatch < h.batch ? nbatch : h.batch;
    int rc = mfft_stage_convert(h.d_data, h.fmt, h.d_f64, MFFT_FMT_C64, inner, nb, nullptr, 0, nullptr, 0, 1.0, 0);
    if (!rc && hipMemcpy(result, h.d_f64, inner * nb * 16, hipMemcpyDeviceToHost) != hipSuccess) rc = MFFT_ERR_HIP;
    set_err(rc);
}

}  // namespace

extern "C" {

int tcfftGetLastError(void) { return g_compat_err; }

void tcfftCreate(tcfftHandle *plan, int n, int n_batch)
{
    if (!plan) {
        set_err(MFFT_ERR_INVALID);
        return;
    }
    memset(plan, 0, sizeof *plan);
    plan->N = n;
    plan->N_batch = n_batch;
    std::vector<int> r = radix_list(n);
    plan->n_radices = (int)r.size();
    for (int i = 0; i < 9; ++i) plan->radices[i] = i < (int)r.size() ? r[i] : 16;
    plan->impl = make_compat(n, 0, n_batch, env_precision(), true);
    if (plan->impl) {
        mfft_info inf;
        mfft_plan_info(((Compat *)plan->impl)->plan, &inf);
        plan->n_mergings = inf.npasses;
    }
}

void tcfftExec(tcfftHandle plan, void *data) { compat_exec((Compat *)plan.impl, data, 0); }

void tcfftDestroy(tcfftHandle *plan)
{
    if (plan && plan->impl) {
        delete (Compat *)plan->impl;
        plan->impl = nullptr;
    }
}

void tcfftCreate_2d(tcfftHandle_2d *plan, int nx, int ny, int n_batch)
{
    if (!plan) {
        set_err(MFFT_ERR_INVALID);
        return;
    }
    memset(plan, 0, sizeof *plan);
    plan->Nx = nx;
    plan->Ny = ny;
    plan->N_batch = n_batch;
    std::vector<int> rx = radix_list(nx), ry = radix_list(ny);
    plan->n_radices_x = (int)rx.size();
    plan->n_radices_y = (int)ry.size();
    for (int i = 0; i < 3; ++i) {
        plan->radices_x[i] = i < (int)rx.size() ? rx[i] : 16;
        plan->radices_y[i] = i < (int)ry.size() ? ry[i] : 16;
    }
    if (rx.size() > 3 || ry.size() > 3) {
        set_err(MFFT_ERR_UNSUPPORTED);
        return;
    }
    plan->impl = make_compat(nx, ny, n_batch, env_precision(), true);
}

void tcfftExec_2d(tcfftHandle_2d plan, void *data) { compat_exec((Compat *)plan.impl, data, 0); }

void tcfftDestroy_2d(tcfftHandle_2d *plan)
{
    if (plan && plan->impl) {
        delete (Compat *)plan->impl;
        plan->impl = nullptr;
    }
}

void setup(double *data, int n, int n_batch) { harness_setup(g_h1, data, n, 0, n_batch, env_precision()); }
void doit(int iter) { harness_doit(g_h1, iter); }
void finalize(double *result, int N_batch) { harness_finalize(g_h1, result, N_batch); }

void setup_2d(double *data, int nx, int ny, int n_batch) { harness_setup(g_h2, data, nx, ny, n_batch, env_precision()); }
void doit_2d(int iter) { harness_doit(g_h2, iter); }
void finalize_2d(double *result, int N_batch) { harness_finalize(g_h2, result, N_batch); }

void setup_cu(double *data, int n, int n_batch) { harness_setup(g_hcu, data, n, 0, n_batch, MFFT_PREC_FP32); }
void doit_cu(int iter) { harness_doit(g_hcu, iter); }
void finalize_cu(double *result, int n, int n_batch)
{
    (void)n;
    harness_finalize(g_hcu, result, n_batch);
}

void setup_cu_2d(double *data, int nx, int ny, int n_batch)
{
    harness_setup(g_hcu2, data, nx, ny, n_batch, MFFT_PREC_FP32);
}
void doit_cu_2d(int iter) { harness_doit(g_hcu2, iter); }
void finalize_cu_2d(double *result, int nx, int ny, int n_batch)
{
    (void)nx;
    (void)ny;
    harness_finalize(g_hcu2, result, n_batch);
}

}  // extern "C"
