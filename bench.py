"""Benchmark of the matrix-core FFT hot path on one MI355X.

Default workload (the metric's configuration): batched 1D complex FFT, N=65536 (radix
16^4, two 256-point passes), batch=256, fp16 interleaved input and output, split-fp16
(hi+lo) matrix-core stages with fp32 accumulation and per-stage 1/16 scaling. One step =
one exec of the whole batch with the input already resident in HBM.

Printed (one JSON line): GFLOP/s = 5*N*log2(N)*batch / t, rel-L2 and rel-max vs the FP64
oracle on a sample of the batch, the plain fp32 GPU FFT (this library's fp32 VALU path)
and the fp16 / bf16 variants on the same sizes, torch.fft (rocFFT, fp32) as a labelled
external data point, the roofline of the dominant kernel from HIP events on the exec
stream, and the FP64 CPU baseline on the host cores.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c2|c3]
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

CONFIGS = {
    # name: (shape, batch, description)
    "c2": ((65536,), 256, "batched 1D complex FFT N=65536 (16^4) x 256, fp16 I/O"),
    "c1": ((256,), 4096, "batched 1D complex FFT N=256 x 4096, fp16 I/O"),
    "c3": ((1024, 1024), 64, "batched 2D complex FFT 1024x1024 x 64, fp16 I/O"),
}


def flops(shape, batch):
    n = math.prod(shape)
    return 5.0 * n * math.log2(n) * batch


def make_input(shape, batch, seed=42):
    import numpy as np

    import oracle

    n = math.prod(shape)
    v = oracle.round_fp16(oracle.uniform(batch * n * 2, seed)).reshape(batch, *shape, 2)
    return v


def time_plan(plan, x, out, steps, warmup, stream):
    """Run `steps` execs with per-launch HIP events; return (total_ms, per-pass avg ms)."""
    import torch

    from fft_demo_amd.hiptime import EventPool

    nl = plan.info["launches_per_exec"]
    npass = len(plan.info["passes"])
    for _ in range(warmup):
        plan.exec(x, out, stream=stream)
    torch.cuda.synchronize()
    pool = EventPool(steps * (nl + 1))
    lib = __import__("fft_demo_amd._lib", fromlist=["load"]).load()
    torch.cuda.synchronize()
    for i in range(steps):
        rc = lib.mfft_exec_events(plan._h, x.data_ptr(), out.data_ptr(), stream.cuda_stream, pool.slice(i * (nl + 1)),
                                  nl + 1)
        if rc != 0:
            raise RuntimeError(lib.mfft_last_error().decode())
    torch.cuda.synchronize()
    total = pool.elapsed_ms(0, steps * (nl + 1) - 1)
    per_pass = [0.0] * npass
    per_pass_n = [0] * npass
    for i in range(steps):
        base = i * (nl + 1)
        for k in range(nl):
            d = pool.elapsed_ms(base + k, base + k + 1)
            per_pass[k % npass] += d
            per_pass_n[k % npass] += 1
    pool.close()
    # wall time of the steps themselves (first launch of step 0 .. end of the last step)
    return total, [per_pass[i] / max(1, per_pass_n[i]) for i in range(npass)]


def accuracy(plan, xh, out, sample, direction=-1):
    import numpy as np

    import oracle

    y = out[:sample].float().cpu().numpy().astype(np.float64)
    yc = (y[..., 0] + 1j * y[..., 1]) / plan.out_scale
    xc = xh[:sample, ..., 0] + 1j * xh[:sample, ..., 1]
    ref = oracle.fft(xc, sign=direction) if len(plan.shape) == 1 else oracle.fft2d(xc, sign=direction)
    return oracle.err_l2max(yc, ref)


def run_variant(name, shape, batch, xh, steps, warmup, dev, stream, sample):
    import torch

    import fft_demo_amd as F

    prec = name
    plan = F.Plan(shape, batch, precision=prec)
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
    (main.cpp:68-80) on 1 core, plus the OpenMP-over-batch variant on all cores."""
    import numpy as np

    import oracle

    n = math.prod(shape)
    sample = batch if len(shape) == 1 else max(1, batch // 16)
    data = np.ascontiguousarray(xh[:sample].reshape(-1))
    runs = []
    for _ in range(3):
        t = time.perf_counter()
        if len(shape) == 1:
            oracle.fft_batch_loop(data, n, sample)
        else:
            xc = xh[:sample, ..., 0] + 1j * xh[:sample, ..., 1]
            oracle.fft2d(xc, threads=1)
        runs.append(time.perf_counter() - t)
    t1 = sorted(runs)[len(runs) // 2]
    ncores = os.cpu_count() or 1
    xc = xh[:sample, ..., 0] + 1j * xh[:sample, ..., 1]
    runs = []
    for _ in range(3):
        t = time.perf_counter()
        if len(shape) == 1:
            oracle.fft(xc, threads=0)
        else:
            oracle.fft2d(xc, threads=0)
        runs.append(time.perf_counter() - t)
    ta = sorted(runs)[len(runs) // 2]
    f = flops(shape, sample)
    return {
        "value": f / t1 / 1e9,
        "unit": "GFLOP/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{sample} transforms of the workload, FP64 Stockham oracle, single-thread batch loop "
                  f"(reference harness shape main.cpp:68-80), median of 3",
        "all_cores": {"value": f / ta / 1e9, "unit": "GFLOP/s", "cores": ncores},
        "host_cpu": _cpu_model(),
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="headline precision only")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the batched FFT shards along "
              f"the batch with no exchange, so the per-GPU figure is the 1-GPU figure.", file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    shape, batch, desc = CONFIGS[args.config]
    xh = make_input(shape, batch)
    sample = 4 if len(shape) == 1 else 1

    head = run_variant("split_fp16", shape, batch, xh, args.steps, args.warmup, dev, stream, sample)
    comps = {}
    if not args.no_compare:
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
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": f"fft_tile_pow2 pass {k} (R={pinfo['R']}, lines/tile={pinfo['lines_per_tile']}, "
                  f"lane_lines={pinfo['lane_lines']})",
        "algorithmic_bytes_per_launch": algo_bytes,
        "avg_launch_ms": round(pp[k], 5),
    }
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
        "data": "synthetic: re/im uniform[-1,1] from SplitMix64 seed 42, rounded to fp16",
        "config": {"workload": desc, "shape": list(shape), "batch": batch, "precision": "split_fp16 (hi+lo, fp32 acc)",
                   "io": "c16 in / c16 out", "scale": "1/16 per radix-16 stage (out = DFT/N)"},
        "rel_l2": head["rel_l2"],
        "rel_max": head["rel_max"],
        "accuracy_sample": f"{sample} of {batch} transforms vs FP64 oracle",
        "whole_exec_hbm_gbs": round(whole_bytes / (head["ms_per_exec"] * 1e-3) / 1e9, 1),
        "passes": head["passes"],
        "pass_avg_ms": [round(v, 5) for v in pp],
        "batch_chunk": head["batch_chunk"],
        "comparators": comps,
        "ratio_vs_plain_fp32": round(head["gflops"] / comps["fp32"]["gflops"], 3) if "fp32" in comps else None,
        "roofline": roofline,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(shape, batch, xh)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
