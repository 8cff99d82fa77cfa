"""Benchmark of the matrix-core FFT hot path on one MI355X.

Default workload (the metric's configuration): batched 1D complex FFT, N=65536 (radix
16^4, two 256-point passes), batch=256, fp16 interleaved input and output, split-fp16
(hi+lo) matrix-core stages with fp32 accumulation and per-stage 1/16 scaling, fp16
intermediates between the two passes (the fp32-intermediate variant is reported beside
it). One step = one exec of the whole batch with the input already resident in HBM.

Printed (one JSON line): GFLOP/s = 5*N*log2(N)*batch / t, rel-L2 and rel-max vs the FP64
oracle on a sample of the batch, the plain fp32 GPU FFT (this library's fp32 VALU path)
and the fp16 / bf16 variants on the same sizes, torch.fft (rocFFT, fp32) as a labelled
external data point, the roofline of the dominant kernel from HIP events on the exec
stream, and the FP64 CPU baseline on the host cores.

Serving (the "serve" object of the line): the continuous-batching request server on the
mixed request workload -- Poisson arrivals at 2000 requests/s (seed 42), 60% 1D FFT,
25% 2D FFT, 15% conv2d against 32 resident filters, 64 GB HBM budget, split fp16 --
replayed in real time: p50/p99 latency and throughput, the same trace through the
server's plain mode (fp32, one request at a time), saturation throughput of both and of
torch.fft (rocFFT) called once per request, and the highest arrival rate each sustains
within one p99 bound. --config c5 runs the full 30 s trace; the default line carries a
shorter one (--serve-seconds).

--config c4 times the FFT 'same' convolution (2048x2048 x 16, 31x31 filter, pad 2176).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c2|c3|c4|c5]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FFT GFLOP/s (5NlogN) at rel-L2≤1e-3 vs plain FP32 GPU FFT; serve p50/p99 ms"
HBM_PEAK_GBS = 8000.0
TUNE = True
TIMING = ""
PASS_TIMING = ""

CONFIGS = {
    # name: (shape, batch, description)
    "c2": ((65536,), 256, "batched 1D complex FFT N=65536 (16^4) x 256, fp16 I/O"),
    "c1": ((256,), 4096, "batched 1D complex FFT N=256 x 4096, fp16 I/O"),
    "c3": ((1024, 1024), 64, "batched 2D complex FFT 1024x1024 x 64, fp16 I/O"),
}


def flops(shape, batch):
    n = math.prod(shape)
    return 5.0 * n * math.log2(n) * batch


def make_input(shape, batch, seed=42, spikes=False):
    """Interleaved re/im uniform[-1,1] (SplitMix64), rounded to fp16. spikes: 10% of the
    complex entries (mask from the next seed) scaled x100, to stress dynamic range."""
    import numpy as np

    import oracle

    n = math.prod(shape)
    v = oracle.uniform(batch * n * 2, seed).reshape(batch * n, 2)
    if spikes:
        mask = oracle.uniform(batch * n, seed + 1, 0.0, 1.0) < 0.1
        v[mask] *= 100.0
    return oracle.round_fp16(v).reshape(batch, *shape, 2)


def round_trip(shape, batch, xh, dev, sample):
    """Forward (per-stage 1/radix) then inverse (unscaled) split-fp16 transforms with fp16
    intermediates on `sample` items: (rel_l2, rel_max) of the round trip against the input."""
    import numpy as np
    import torch

    import fft_demo_amd as F
    import oracle

    fwd = F.Plan(shape, sample, precision="split_fp16", mid_fmt="c16")
    inv = F.Plan(shape, sample, precision="split_fp16", mid_fmt="c16", scale=0, direction=1)
    x = torch.from_numpy(np.ascontiguousarray(xh[:sample])).to(device=dev, dtype=fwd.in_dtype()).contiguous()
    back = inv.exec(fwd.exec(x)).float().cpu().numpy().astype(np.float64)
    torch.cuda.synchronize()
    l2, mx = oracle.err_l2max(back[..., 0] + 1j * back[..., 1], xh[:sample, ..., 0] + 1j * xh[:sample, ..., 1])
    fwd.close()
    inv.close()
    return l2, mx


def time_plan(plan, x, out, steps, warmup, stream):
    """Time `steps` back-to-back execs between two HIP events on the exec stream; when an
    exec is shorter than 0.5 ms (the Python wrapper's launch overhead would then leave gaps
    in the queue) a hipGraph holding the `steps` execs is timed too and the faster is reported.
    Each pass is then timed (for the roofline): `steps` whole execs run pass by pass
    (mfft_exec_pass), the pass's launch in each between a start and a stop HIP event that
    the runtime records around the kernel itself on the exec stream (its execution window,
    as a kernel trace reports it); the average is the pass's launch time. (A pass launched
    alone over and over would find the input it rereads in the last-level cache.) Plans that run their batch in
    several chunks fall back to two events around `steps` back-to-back runs of the pass
    (one launch per chunk, launch gaps included), divided by the number of launches.
    Returns (total_ms, [per-pass avg ms per launch])."""
    import torch

    from fft_demo_amd.hiptime import EventPool

    npass = len(plan.info["passes"])
    for _ in range(warmup):
        plan.exec(x, out, stream=stream)
    torch.cuda.synchronize()
    lib = __import__("fft_demo_amd._lib", fromlist=["load"]).load()
    ends = EventPool(2)
    ends.record(0, stream.cuda_stream)
    for _ in range(steps):
        plan.exec(x, out, stream=stream)
    ends.record(1, stream.cuda_stream)
    torch.cuda.synchronize()
    total = ends.elapsed_ms(0, 1)
    global TIMING
    TIMING = "back-to-back execs between two HIP events"
    if total / steps < 0.5:
        # short execs (host launch overhead of the Python wrapper is comparable to the GPU
        # time): also capture the steps in one hipGraph and time its replay; the faster of
        # the two ways of running the same execs is reported (graph nodes carry a launch
        # latency of their own, so neither wins everywhere)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cs = torch.cuda.current_stream()  # the capture stream
            for _ in range(steps):
                plan.exec(x, out, stream=cs)
        g.replay()
        torch.cuda.synchronize()
        ends.record(0, stream.cuda_stream)
        g.replay()
        ends.record(1, stream.cuda_stream)
        torch.cuda.synchronize()
        tg = ends.elapsed_ms(0, 1)
        if tg < total:
            total = tg
            TIMING = "hipGraph replay of the steps (captured execs), between two HIP events"
        del g
    ends.close()
    global PASS_TIMING
    per_pass = []
    single = plan.info["batch_chunk"] >= plan.batch
    nchunks = -(-plan.batch // plan.info["batch_chunk"])
    PASS_TIMING = ("kernel window: start/stop HIP events recorded around the pass's launch "
                   "(hipExtLaunchKernel) inside whole execs run pass by pass" if single else
                   f"two HIP events around back-to-back runs of the pass ({nchunks} chunk launches each), per launch")
    pool = EventPool(2 * steps)
    for k in range(npass):
        if single:
            # every exec runs all its passes in order (so the caches hold what they hold in
            # an exec); only pass k's launch carries the events
            for i in range(steps):
                for j in range(npass):
                    ev = (pool.arr[2 * i], pool.arr[2 * i + 1]) if j == k else (None, None)
                    rc = lib.mfft_exec_pass(plan._h, j, x.data_ptr(), out.data_ptr(), stream.cuda_stream, *ev)
                    if rc != 0:
                        raise RuntimeError(lib.mfft_last_error().decode())
            torch.cuda.synchronize()
            per_pass.append(sum(pool.elapsed_ms(2 * i, 2 * i + 1) for i in range(steps)) / steps)
        else:
            pool.record(0, stream.cuda_stream)
            for _ in range(steps):
                rc = lib.mfft_exec_pass(plan._h, k, x.data_ptr(), out.data_ptr(), stream.cuda_stream, None, None)
                if rc != 0:
                    raise RuntimeError(lib.mfft_last_error().decode())
            pool.record(1, stream.cuda_stream)
            torch.cuda.synchronize()
            per_pass.append(pool.elapsed_ms(0, 1) / steps / nchunks)
    pool.close()
    # leave `out` holding a full exec's result
    plan.exec(x, out, stream=stream)
    torch.cuda.synchronize()
    return total, per_pass


def accuracy(plan, xh, out, sample, direction=-1):
    import numpy as np

    import oracle

    y = out[:sample].float().cpu().numpy().astype(np.float64)
    yc = (y[..., 0] + 1j * y[..., 1]) / plan.out_scale
    xc = xh[:sample, ..., 0] + 1j * xh[:sample, ..., 1]
    ref = oracle.fft(xc, sign=direction) if len(plan.shape) == 1 else oracle.fft2d(xc, sign=direction)
    return oracle.err_l2max(yc, ref)


def run_variant(name, shape, batch, xh, steps, warmup, dev, stream, sample, mid_fmt=None):
    import torch

    import fft_demo_amd as F
    from fft_demo_amd import tune

    prec = name
    # matrix-core variants scale 1/radix per stage (fp16/bf16 would overflow on the 2D
    # workload's x100 entries otherwise); plain fp32 is unscaled
    scale = 0 if prec == "fp32" else 1
    # every variant gets the same treatment: autotuned schedule (split point, batch chunk),
    # unless --no-tune (profiling runs: only the planner's schedule is launched)
    sched = tune.autotune(shape, batch, precision=prec, mid_fmt=mid_fmt, scale=scale) if TUNE else None
    plan = F.Plan(shape, batch, precision=prec, split_hint=tune.SPLIT_TUNED if TUNE else 0, mid_fmt=mid_fmt,
                  scale=scale)
    dt = plan.in_dtype()
    x = torch.from_numpy(xh).to(device=dev, dtype=dt).contiguous()
    out = plan.empty_out(device=dev)
    total_ms, per_pass = time_plan(plan, x, out, steps, warmup, stream)
    ms = total_ms / steps
    l2, mx = accuracy(plan, xh, out, sample)
    res = {
        "gflops": flops(shape, batch) / (ms * 1e-3) / 1e9,
        "ms_per_exec": ms,
        "rel_l2": l2,
        "rel_max": mx,
        "passes": plan.info["passes"],
        "pass_avg_ms": per_pass,
        "batch_chunk": plan.info["batch_chunk"],
        "launches_per_exec": plan.info["launches_per_exec"],
        "schedule": sched,
        "timing": TIMING,
        "pass_timing": PASS_TIMING,
    }
    del x, out
    plan.close()
    torch.cuda.synchronize()
    return res


def rocfft_point(shape, batch, xh, steps, warmup, dev):
    """torch.fft (hipFFT/rocFFT, complex64) on the same data -- external data point only."""
    import torch

    x = torch.from_numpy(xh).to(device=dev, dtype=torch.float32).contiguous()
    xc = torch.view_as_complex(x)
    dims = tuple(range(1, 1 + len(shape)))
    for _ in range(warmup):
        y = torch.fft.fftn(xc, dim=dims)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        y = torch.fft.fftn(xc, dim=dims)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / steps
    del x, xc, y
    return {"gflops": flops(shape, batch) / (ms * 1e-3) / 1e9, "ms_per_exec": ms, "dtype": "complex64",
            "note": "torch.fft.fftn (rocFFT) - labelled external data point, not called by the library"}


def cpu_baseline(shape, batch, xh):
    """FP64 oracle FFT on the host: the reference's single-thread batch-loop harness shape
    (main.cpp:68-80) on 1 core, plus the OpenMP-over-batch variant on all cores; one
    warm-up run, then the median of five. The reference's own CPU path calls FFTW
    (FFTW_ESTIMATE); libfftw3 is looked up and its absence stated."""
    import ctypes.util

    import numpy as np

    import oracle

    n = math.prod(shape)
    sample = batch if len(shape) == 1 else max(1, batch // 16)
    data = np.ascontiguousarray(xh[:sample].reshape(-1))
    runs = []
    for _ in range(6):
        t = time.perf_counter()
        if len(shape) == 1:
            oracle.fft_batch_loop(data, n, sample)
        else:
            xc = xh[:sample, ..., 0] + 1j * xh[:sample, ..., 1]
            oracle.fft2d(xc, threads=1)
        runs.append(time.perf_counter() - t)
    runs = runs[1:]
    t1 = sorted(runs)[len(runs) // 2]
    ncores = os.cpu_count() or 1
    xc = xh[:sample, ..., 0] + 1j * xh[:sample, ..., 1]
    runs = []
    for _ in range(6):
        t = time.perf_counter()
        if len(shape) == 1:
            oracle.fft(xc, threads=0)
        else:
            oracle.fft2d(xc, threads=0)
        runs.append(time.perf_counter() - t)
    runs = runs[1:]
    ta = sorted(runs)[len(runs) // 2]
    f = flops(shape, sample)
    fftw = ctypes.util.find_library("fftw3")
    return {
        "value": f / t1 / 1e9,
        "unit": "GFLOP/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{sample} transforms of the workload, FP64 Stockham oracle, single-thread batch loop "
                  f"(reference harness shape main.cpp:68-80), one warm-up then median of 5",
        "all_cores": {"value": f / ta / 1e9, "unit": "GFLOP/s", "cores": ncores},
        "host_cpu": _cpu_model(),
        "fftw": f"libfftw3 found ({fftw}), not timed" if fftw else "FFTW unavailable (no libfftw3 on the host)",
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def copy_bandwidth(mib=1024, reps=10):
    """Practical HBM ceiling on this device: a device-to-device copy of `mib` MiB, read +
    write bytes per second (GB/s)."""
    import torch

    a = torch.ones((mib << 20) // 4, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        b.copy_(a)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / reps
    del a, b
    torch.cuda.empty_cache()
    return round(2 * (mib << 20) / (ms * 1e-3) / 1e9, 1)


def profiled_traffic():
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a copy
    kernel of known size), from the summaries committed under profiles/: {key: bytes}, key =
    "<kernel family>:R<points per line>:c<in bits>>c<out bits>"."""
    out = {}
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return out
    for name in sorted(os.listdir(d)):
        if not name.endswith(".md"):
            continue
        with open(os.path.join(d, name)) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    for k, v in json.loads(line.split(":", 1)[1]).items():
                        out[k] = (int(v), name)
    return out


# ------------------------------------------------------------------------- serving
def _rocfft_plain(trace, filters, dev):
    """torch.fft (rocFFT, complex64) once per request, each finished before the next:
    plain per-request FP32 GPU FFT launches. conv2d = irfft2(rfft2(image) * resident
    filter spectrum) over the same padded size as the server's operator."""
    import torch

    from fft_demo_amd import serve
    from fft_demo_amd.conv import good_size

    g = torch.Generator(device="cpu").manual_seed(7)
    maxp = max([serve.points(r) for r in trace if r[1] != serve.KIND_CONV] + [1])
    xin = torch.view_as_complex((torch.rand(maxp, 2, generator=g) * 2 - 1).to(dev))
    side = max(serve.SIDES_CONV)
    img = torch.rand(side * side, generator=g).to(dev)
    fmax = max(f.shape[0] for f in filters)
    spec = {}

    def filt_spec(s, fid):
        key = (s, fid)
        if key not in spec:
            p = good_size(s + fmax - 1)
            f = filters[fid]
            fr = torch.zeros(p, p, device=dev)
            o = (fmax - f.shape[0]) // 2  # centred in the common frame, as in the server
            fr[o:o + f.shape[0], o:o + f.shape[1]] = torch.from_numpy(f).to(dev)
            spec[key] = (p, torch.fft.rfft2(fr))
        return spec[key]

    def one(r):
        _, kind, n, nx, ny, b, fid = r
        if kind == serve.KIND_1D:
            return torch.fft.fft(xin[: n * b].view(b, n))
        if kind == serve.KIND_2D:
            return torch.fft.fft2(xin[: nx * ny * b].view(b, nx, ny))
        p, H = filt_spec(nx, fid)
        c = (fmax - 1) // 2
        y = torch.fft.irfft2(torch.fft.rfft2(img[: nx * ny].view(nx, ny), s=(p, p)) * H, s=(p, p))
        return y[c:c + nx, c:c + ny]

    for r in trace:  # every plan (rocFFT's plan cache) and filter spectrum outside the timed region
        one(r)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for r in trace:
        t = time.perf_counter()
        one(r)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t) * 1e3)
    span = time.perf_counter() - t0
    import numpy as np

    return {"requests": len(trace), "req_per_s": len(trace) / span,
            "gflops": sum(serve.flops(r) for r in trace) / span / 1e9,
            "service_p50_ms": float(np.percentile(lat, 50)), "service_p99_ms": float(np.percentile(lat, 99)),
            "note": "torch.fft (rocFFT, complex64), one request at a time; labelled external baseline"}


def _sustained(make, filters, shapes_trace, target_p99_ms, seconds, start_rate, log):
    """Highest arrival rate (x1.5 steps from start_rate, then four bisection steps) whose
    p99 stays within the bound;
    returns the replay result at that rate. shapes_trace covers the request shapes (plans
    and operators are built from it before the probes)."""
    from fft_demo_amd import serve

    best = None
    rate = start_rate
    S = make()
    serve.register_filters(S, filters)
    ar = serve.Arena(S, shapes_trace, 16 << 30)
    serve.warm(S, shapes_trace, ar)
    def probe(rate):
        r = serve.replay(S, serve.make_trace(rate, seconds, 43), ar)
        log.append({"rate": round(rate, 1), "p99_ms": round(r["p99_ms"], 3), "req_per_s": round(r["req_per_s"], 1),
                    "submit_lag_max_ms": round(r["submit_lag_max_ms"], 3)})
        return r, r["p99_ms"] <= target_p99_ms and r["req_per_s"] >= 0.9 * rate

    try:
        fail = None
        for _ in range(10):
            r, ok = probe(rate)
            if not ok:
                fail = rate
                break
            best, lo = r, rate
            rate *= 1.5
        # four geometric bisection steps between the last rate that held and the first that
        # did not (resolves the knee to about 2.5%)
        if best is not None and fail is not None:
            for _ in range(4):
                mid = (lo * fail) ** 0.5
                r, ok = probe(mid)
                if ok:
                    best, lo = r, mid
                else:
                    fail = mid
    finally:
        S.close()
        del ar
    return best


def serve_section(seconds, sat_requests=2000, p99_bound_ms=2.0):
    import torch

    from fft_demo_amd import serve

    filters = serve.make_filters()
    trace = serve.make_trace(2000.0, seconds, 42)
    out = {"workload": f"Poisson 2000 req/s for {seconds:g} s (seed 42), 60% 1D / 25% 2D / 15% conv2d, "
                       f"32 resident filters, 64 GB HBM budget", "requests": len(trace)}
    modes = {"split": dict(precision="split", mid_fmt="c16"), "plain_fp32": dict(plain=True)}
    for name, kw in modes.items():
        S = serve.Server(**kw)
        serve.register_filters(S, filters)
        ar = serve.Arena(S, trace, 16 << 30)
        serve.warm(S, trace, ar)
        torch.cuda.synchronize()
        r = serve.replay(S, trace, ar)
        sat = serve.replay(S, trace[:sat_requests], ar, time_scale=1e9)
        st = S.stats()
        S.close()
        del ar
        torch.cuda.empty_cache()
        out[name] = {"p50_ms": round(r["p50_ms"], 4), "p99_ms": round(r["p99_ms"], 4),
                     "submit_lag_max_ms": round(r["submit_lag_max_ms"], 3), "scheduler_busy_s": st["busy_us"] / 1e6,
                     "req_per_s": round(r["req_per_s"], 1), "gflops": round(r["gflops"], 1),
                     "saturation_req_per_s": round(sat["req_per_s"], 1),
                     "saturation_gflops": round(sat["gflops"], 1),
                     "launches": st["launches"], "coalesced": st["coalesced"], "paired": st["paired"],
                     "static_bytes": st["static_bytes"], "peak_bytes": st["peak_bytes"]}
    try:
        out["plain_rocfft_external"] = _rocfft_plain(trace[:sat_requests // 2], filters, torch.device("cuda:0"))
    except Exception as e:  # noqa: BLE001
        out["plain_rocfft_external"] = {"error": str(e)}
    torch.cuda.empty_cache()
    # equal p99: the highest arrival rate each mode sustains with p99 within one bound
    bound = p99_bound_ms
    logs = {}
    sus = {}
    for name, kw in modes.items():
        logs[name] = []
        b = _sustained(lambda kw=kw: serve.Server(**kw), filters, trace, bound, min(seconds, 2.0), 2000.0, logs[name])
        sus[name] = round(b["req_per_s"], 1) if b else None
        torch.cuda.empty_cache()
    out["equal_p99"] = {"p99_bound_ms": bound, "sustained_req_per_s": sus, "probes": logs,
                        "ratio": round(sus["split"] / sus["plain_fp32"], 3) if sus["split"] and sus["plain_fp32"]
                        else None}
    out["saturation_ratio_vs_plain_fp32"] = round(out["split"]["saturation_req_per_s"] /
                                                  out["plain_fp32"]["saturation_req_per_s"], 3)
    rf = out["plain_rocfft_external"]
    if "req_per_s" in rf:
        out["saturation_ratio_vs_plain_rocfft"] = round(out["split"]["saturation_req_per_s"] / rf["req_per_s"], 3)
    return out


# ------------------------------------------------------------------------- MFMA use
MFMA_FLOPS = 16 * 16 * 32 * 2  # one v_mfma_f32_16x16x32_f16
MFMA_PEAK_TFLOPS = 2500.0       # MI355X dense fp16


def mfma_utilisation(passes, pass_ms):
    """Matrix-core use of the split-fp16 passes. Every radix-16 stage is a complex DFT16
    on 16 lines at a time: two v_mfma_f32_16x16x32_f16 per real product (the complex
    product is folded into K = 32, the accumulator holds 128 of the 256 complex outputs),
    three products (Fl*Xh + Fh*Xl + Fh*Xh), two for a first stage that reads exact fp16
    input (no Xl). The count per launch agrees with rocprofv3's SQ_INSTS_MFMA
    (profiles/fft_fast_split_n65536x256_mfma.md)."""
    out = []
    for p, ms in zip(passes, pass_ms):
        if p["generic"] or not p["kernel"].startswith("fast"):
            continue
        n16 = sum(1 for r in p["radix"] if r == 16)
        groups = p["points"] // 256
        exact_first = p["in_elem_bytes"] == 4  # c16 in (the split path's fp16 buffers)
        per_group = 2 * (2 if exact_first else 3) + 2 * 3 * (n16 - 1) if n16 else 0
        inst = groups * per_group
        tf = inst * MFMA_FLOPS / (ms * 1e-3) / 1e12
        out.append({"kernel": p["kernel"], "R": p["R"], "instructions_per_launch": inst,
                    "avg_launch_ms": round(ms, 5), "achieved_tflops": round(tf, 1),
                    "peak_tflops": MFMA_PEAK_TFLOPS, "frac": round(tf / MFMA_PEAK_TFLOPS, 4)})
    return out


# ------------------------------------------------------------------------- conv2d
def conv_section(steps, warmup):
    """FFT 'same' convolution: 2048x2048 fp32 images U[0,1], batch 16, one 31x31 filter
    (Gaussian sigma 5 normalised to sum 1, times 1 + U[-0.01, 0.01]), padded to 2176;
    split-fp16 transforms with fp16 image spectra and intermediates and the fp32 filter
    spectrum resident (comparators: split fp16 with fp32 spectra, plain fp32). Accuracy on
    sampled rows of two images against the direct FP64 convolution."""
    import numpy as np
    import torch

    import oracle
    from fft_demo_amd.conv import Conv2d

    B, H, W, K = 16, 2048, 2048, 31
    img = oracle.uniform(B * H * W, 42, 0.0, 1.0).astype(np.float32).reshape(B, 1, H, W)
    c = (K - 1) / 2
    yy, xx = np.mgrid[0:K, 0:K]
    g = np.exp(-((yy - c) ** 2 + (xx - c) ** 2) / (2 * 25.0))
    g /= g.sum()
    g *= 1 + oracle.uniform(K * K, 43, -0.01, 0.01).reshape(K, K)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(img).to(dev)
    out = torch.empty((B, 1, H, W), dtype=torch.float32, device=dev)
    gd = torch.from_numpy(g.astype(np.float32)).to(dev).contiguous()

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(steps):
            fn()
        e.record()
        torch.cuda.synchronize()
        return s.elapsed_time(e) / steps

    def accuracy():
        errs = []
        for b in (0, B - 1):
            ref = oracle.conv2d_same(img[b, 0].astype(np.float64), g, rows=rows)
            got = out[b, 0, rows[0]:rows[1]].double().cpu().numpy()
            errs.append((float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
                         float(np.abs(got - ref).max() / np.abs(ref).max())))
        return max(r[0] for r in errs), max(r[1] for r in errs)

    rows = (0, 48)
    res = {}
    for name, prec, spec16 in (("split_fp16", "split_fp16", True), ("split_fp16_fp32_spectra", "split_fp16", False),
                               ("fp32", "fp32", False)):
        op = Conv2d(B, 1, 1, H, W, K, K, mode="same", precision=prec, spec_fp16=spec16)
        op.set_filters(gd)
        ms = timed(lambda: op.forward(x, out))
        l2, mx = accuracy()
        res[name] = {"images_per_s": round(B / (ms * 1e-3), 1), "ms_per_batch": round(ms, 4), "rel_l2": l2,
                     "rel_max": mx, "pad": list(op.pad), "spectra": "fp16" if spec16 else "fp32"}
        P = op.pad[0] * op.pad[1]
        op.close()
    # labelled external point: torch.fft real transforms (rocFFT), resident filter spectrum
    p0, p1 = res["split_fp16"]["pad"]
    fr = torch.zeros(p0, p1, device=dev)
    fr[:K, :K] = gd
    H_spec = torch.fft.rfft2(fr)
    c0 = (K - 1) // 2

    def torch_conv():
        y = torch.fft.irfft2(torch.fft.rfft2(x[:, 0], s=(p0, p1)) * H_spec, s=(p0, p1))
        out[:, 0] = y[:, c0:c0 + H, c0:c0 + W]

    ms = timed(torch_conv)
    l2, mx = accuracy()
    res["rocfft_fp32_external"] = {"images_per_s": round(B / (ms * 1e-3), 1), "ms_per_batch": round(ms, 4),
                                   "rel_l2": l2, "rel_max": mx,
                                   "note": "torch.fft rfft2/irfft2 (rocFFT) with the filter spectrum precomputed"}
    head = res["split_fp16"]
    ms = head["ms_per_batch"]
    return {"workload": "FFT 'same' conv2d, 2048x2048 fp32 x 16, 31x31 Gaussian filter, pad 2176, split fp16, "
                        "fp16 spectra",
            "images_per_s": head["images_per_s"], "ms_per_batch": ms,
            "nominal_gflops": round(B * 2 * 5.0 * P * math.log2(P) / (ms * 1e-3) / 1e9, 1),
            "pad": head["pad"], "rel_l2": head["rel_l2"], "rel_max": head["rel_max"],
            "accuracy_sample": f"rows {rows[0]}..{rows[1] - 1} of images 0 and {B - 1} vs direct FP64 convolution",
            "io_gbs": round(B * H * W * 4 * 2 / (ms * 1e-3) / 1e9, 1),
            "comparators": {"split_fp16_fp32_spectra": res["split_fp16_fp32_spectra"], "fp32": res["fp32"],
                            "rocfft_fp32_external": res["rocfft_fp32_external"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 200 execs of 65536 x 256 take ~11 ms: long enough for the clocks and caches to
    # settle (20 read a few percent slow, each pass's launches included); the 256 x 4096
    # workload (~4 us an exec) times 1000 execs in one graph by default
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 200; 1000 for c1)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS) + ["c4", "c5"])
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="trace length of the serving run (c5: 30)")
    ap.add_argument("--serve-p99-ms", type=float, default=2.0, help="p99 bound of the equal-p99 comparison")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="planner's schedule, no autotuning (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (the known-size kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="headline precision only")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 1000 if args.config == "c1" else 200
    global TUNE
    TUNE = not args.no_tune
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the batched FFT shards along "
              f"the batch with no exchange, so the per-GPU figure is the 1-GPU figure.", file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                          "launches": 5}))
        return
    if args.config == "c4":
        cv = conv_section(args.steps, args.warmup)
        print(json.dumps({"metric": METRIC, "value": cv["images_per_s"], "unit": "images/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": cv["ms_per_batch"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp16",
                          "data": "synthetic", "config": {"workload": cv["workload"]}, "conv": cv}))
        return
    if args.config == "c5":
        sv = serve_section(30.0, p99_bound_ms=args.serve_p99_ms)
        print(json.dumps({"metric": METRIC, "value": sv["split"]["p99_ms"], "unit": "ms (p99)", "n_gpus": 1,
                          "steps": sv["requests"], "warmup": 0, "ms_per_step": None, "higher_is_better": False,
                          "scaling": "weak", "vs_baseline": None, "dtype": "fp16", "data": "synthetic",
                          "config": {"workload": sv["workload"]}, "serve": sv}))
        return
    shape, batch, desc = CONFIGS[args.config]
    spikes = args.config == "c3"  # the 2D workload carries 10% entries x100
    xh = make_input(shape, batch, spikes=spikes)
    sample = 4 if len(shape) == 1 else 1

    # headline: split fp16 with fp16 intermediates between passes (one rounding per pass
    # boundary; rel-L2 stays within 1e-3); the fp32-intermediate variant is a comparator
    head = run_variant("split_fp16", shape, batch, xh, args.steps, args.warmup, dev, stream, sample, mid_fmt="c16")
    comps = {}
    if not args.no_compare:
        comps["split_fp16_c32_intermediates"] = run_variant("split_fp16", shape, batch, xh, args.steps, args.warmup,
                                                            dev, stream, sample)
        for name in ("fp32", "fp16", "bf16"):
            comps[name] = run_variant(name, shape, batch, xh, args.steps, args.warmup, dev, stream, sample)
        try:
            comps["rocfft_fp32_external"] = rocfft_point(shape, batch, xh, args.steps, args.warmup, dev)
        except Exception as e:  # noqa: BLE001
            comps["rocfft_fp32_external"] = {"error": str(e)}
    for v in comps.values():
        v.pop("passes", None)

    # roofline of the dominant kernel (the pass with the largest average launch time)
    pp = head["pass_avg_ms"]
    k = max(range(len(pp)), key=lambda i: pp[i])
    pinfo = head["passes"][k]
    algo_bytes = pinfo["points"] * (pinfo["in_elem_bytes"] + pinfo["out_elem_bytes"])
    achieved = algo_bytes / (pp[k] * 1e-3) / 1e9
    tkey = f"{pinfo['kernel']}:R{pinfo['R']}:c{pinfo['in_elem_bytes'] * 4}>c{pinfo['out_elem_bytes'] * 4}"
    tmeas = profiled_traffic().get(tkey)
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": tmeas[0] if tmeas else None,
        "traffic_source": f"profiles/{tmeas[1]} ({tkey})" if tmeas else None,
        "kernel": f"pass {k}: {pinfo['kernel']} (R={pinfo['R']}, lines/tile={pinfo['lines_per_tile']}, "
                  f"c{pinfo['in_elem_bytes'] * 4} in / c{pinfo['out_elem_bytes'] * 4} out)",
        "algorithmic_bytes_per_launch": algo_bytes,
        "avg_launch_ms": round(pp[k], 5),
    }
    roofline["copy_gbs"] = copy_bandwidth()
    roofline["frac_of_copy"] = round(achieved / roofline["copy_gbs"], 4)
    mfma = mfma_utilisation(head["passes"], pp)
    n = math.prod(shape)
    whole_bytes = batch * n * 8  # fp16 complex in + out
    line = {
        "metric": METRIC,
        "value": round(head["gflops"], 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_exec"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16",
        "data": "synthetic: re/im uniform[-1,1] from SplitMix64 seed 42" +
                (", 10% of entries x100 (mask seed 43)" if spikes else "") + ", rounded to fp16",
        "config": {"workload": desc, "shape": list(shape), "batch": batch, "precision": "split_fp16 (hi+lo, fp32 acc)",
                   "io": "c16 in / c16 out", "intermediates": "c16 between passes",
                   "scale": "1/16 per radix-16 stage (out = DFT/N)"},
        "rel_l2": head["rel_l2"],
        "rel_max": head["rel_max"],
        "accuracy_sample": f"{sample} of {batch} transforms vs FP64 oracle",
        "whole_exec_hbm_gbs": round(whole_bytes / (head["ms_per_exec"] * 1e-3) / 1e9, 1),
        "passes": head["passes"],
        "pass_avg_ms": [round(v, 5) for v in pp],
        "pass_timing": head["pass_timing"],
        "batch_chunk": head["batch_chunk"],
        "schedule": head["schedule"],
        "timing": head["timing"],
        "comparators": comps,
        "ratio_vs_plain_fp32": round(head["gflops"] / comps["fp32"]["gflops"], 3) if "fp32" in comps else None,
        "roofline": roofline,
        "mfma": mfma,
    }
    if len(shape) == 2:
        rl2, rmx = round_trip(shape, batch, xh, dev, sample)
        line["round_trip"] = {"rel_l2": rl2, "rel_max": rmx,
                              "sample": f"{sample} transform(s): forward (1/radix per stage) then inverse, vs the input"}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(shape, batch, xh)
    if not args.no_serve:
        torch.cuda.empty_cache()
        line["serve"] = serve_section(args.serve_seconds, p99_bound_ms=args.serve_p99_ms)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
