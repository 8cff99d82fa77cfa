"""Host side of the continuous-batching request server (C ABI: mfft_server_* in include/mfft.h).

The reference has no serving layer; its nearest idea is the multi-stream batch split of
fft_2d_c2c (tensorcore_fft_conv2d.cpp:79-146). This module

  * wraps the server (Server: submit / poll / wait_idle / stats),
  * generates the mixed request workload: Poisson arrivals, 60% 1D FFT (N = 2^k,
    k ~ U{8..20}, batch ~ U{1..64}), 25% 2D FFT (side in {256, 512, 1024, 2048},
    batch ~ U{1..16}), 15% conv2d ('same', image side in {512, 768, 1024, 1536, 2048},
    one of 32 resident filters of odd size 3..63, Gaussian-weighted, sigma = 5,
    normalised to sum 1, plus +-1% uniform noise),
  * replays a trace against a server in real time and reports latency percentiles and
    throughput.

Device buffers come from torch (allocation only); all transforms run in the HIP library.
"""
import collections
import ctypes
import gc
import time

import numpy as np

from . import _lib

KIND_1D, KIND_2D, KIND_CONV = _lib.REQ_FFT1D, _lib.REQ_FFT2D, _lib.REQ_CONV2D
SIDES_2D = (256, 512, 1024, 2048)
SIDES_CONV = (512, 768, 1024, 1536, 2048)
N_FILTERS = 32


class ServerError(RuntimeError):
    pass


# One finished request (mfft_completion copied out of the poll buffer).
Completion = collections.namedtuple("Completion", "tag status submit_us start_us end_us out_scale")


def _check(lib, rc, what):
    if rc != 0:
        raise ServerError(f"{what}: rc={rc}: {lib.mfft_server_last_error().decode()}")


class Server:
    """A running request server on the current device.

    precision: "split" (split fp16, default), "fp16", "bf16" or "fp32"; plain=True gives
    the baseline (fp32, one request at a time, each finished before the next starts).
    mid_fmt="c16" keeps the split plans' intermediates, and the conv operators' image
    spectra, in fp16 (one more rounding per pass boundary, half the intermediate traffic).
    FFT requests carry interleaved complex data in the server's format (fp16 pairs for
    split/fp16, bf16 pairs for bf16, fp32 pairs for fp32 and plain) and produce
    out_scale * DFT(in); conv2d requests carry fp32 real images.
    """

    PREC = {"fp16": 0, "bf16": 1, "split": 2, "fp32": 3}

    def __init__(self, precision="split", plain=False, hbm_budget=64 << 30, streams=4, max_batch_1d=64,
                 max_batch_2d=16, coalesce_max=256, slab_bytes=512 << 20, mid_fmt=None):
        self.L = _lib
        self.lib = _lib.load()
        d = _lib.ServerDesc()
        self.lib.mfft_server_desc_default(ctypes.byref(d))
        d.hbm_budget = int(hbm_budget)
        d.streams = int(streams)
        d.precision = self.PREC[precision]
        d.max_batch_1d = int(max_batch_1d)
        d.max_batch_2d = int(max_batch_2d)
        d.coalesce_max = int(coalesce_max)
        d.slab_bytes = int(slab_bytes)
        d.plain = 1 if plain else 0
        d.mid_fmt = -1 if mid_fmt is None else {"c16": 0, "cbf16": 1, "c32": 2}[mid_fmt]
        self.desc = d
        self.plain = plain
        self.precision = "fp32" if plain else precision
        self.handle = ctypes.c_void_p()
        self._poll_buf = None
        _check(self.lib, self.lib.mfft_server_create(ctypes.byref(self.handle), ctypes.byref(d)), "create")

    # -- data format -------------------------------------------------------------------
    @property
    def torch_dtype(self):
        import torch

        return {"split": torch.float16, "fp16": torch.float16, "bf16": torch.bfloat16, "fp32": torch.float32}[
            self.precision]

    def add_filter(self, fid, weights):
        w = np.ascontiguousarray(weights, dtype=np.float32)
        assert w.ndim == 2
        _check(self.lib, self.lib.mfft_server_add_filter(self.handle, int(fid), w.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float)), w.shape[0], w.shape[1]), "add_filter")

    def submit(self, kind, inp_ptr, out_ptr, tag, n=0, nx=0, ny=0, batch=1, filter_id=0, direction=-1):
        r = _lib.Request(kind=kind, direction=direction, n=n, nx=nx, ny=ny, batch=batch, in_=inp_ptr, out=out_ptr,
                         filter_id=filter_id, tag=tag)
        _check(self.lib, self.lib.mfft_server_submit(self.handle, ctypes.byref(r)), "submit")

    def poll(self, max_items=4096):
        """Up to max_items finished requests as Completion tuples. The ctypes buffer is
        allocated once per size and reused; the records are copied out of it, so a caller
        that keeps them holds no buffer."""
        buf = self._poll_buf
        if buf is None or len(buf) < max_items:
            buf = self._poll_buf = (_lib.Completion * max_items)()
        n = ctypes.c_int(0)
        _check(self.lib, self.lib.mfft_server_poll(self.handle, buf, max_items, ctypes.byref(n)), "poll")
        return [Completion(c.tag, c.status, c.submit_us, c.start_us, c.end_us, c.out_scale)
                for c in buf[:n.value]]

    def wait_idle(self, timeout_s=120.0):
        _check(self.lib, self.lib.mfft_server_wait_idle(self.handle, float(timeout_s)), "wait_idle")

    def stats(self):
        st = _lib.ServerStats()
        _check(self.lib, self.lib.mfft_server_get_stats(self.handle, ctypes.byref(st)), "stats")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.handle:
            self.lib.mfft_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# -- workload ---------------------------------------------------------------------------
def make_filters(seed=42, count=N_FILTERS):
    """The resident filters: odd sizes 3..63, Gaussian weights (sigma 5) normalised to sum
    1, times (1 + u), u ~ U[-0.01, 0.01]."""
    rng = np.random.default_rng(seed)
    sizes = [3 + 2 * (i % 31) for i in range(count)]
    out = []
    for s in sizes:
        c = (s - 1) / 2
        y, x = np.mgrid[0:s, 0:s]
        g = np.exp(-((y - c) ** 2 + (x - c) ** 2) / (2 * 5.0 ** 2))
        g /= g.sum()
        g *= 1.0 + rng.uniform(-0.01, 0.01, size=g.shape)
        out.append(g.astype(np.float32))
    return out


def make_trace(rate=2000.0, duration=30.0, seed=42, max_requests=None):
    """Poisson arrivals; each entry (t_arrival_s, kind, n, nx, ny, batch, filter_id)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration or (max_requests is not None and len(out) >= max_requests):
            break
        u = rng.random()
        if u < 0.60:
            out.append((t, KIND_1D, 1 << int(rng.integers(8, 21)), 0, 0, int(rng.integers(1, 65)), 0))
        elif u < 0.85:
            s = SIDES_2D[int(rng.integers(0, 4))]
            out.append((t, KIND_2D, 0, s, s, int(rng.integers(1, 17)), 0))
        else:
            s = SIDES_CONV[int(rng.integers(0, len(SIDES_CONV)))]
            out.append((t, KIND_CONV, 0, s, s, 1, int(rng.integers(0, N_FILTERS))))
    return out


def points(req):
    _, kind, n, nx, ny, batch, _ = req
    return n * batch if kind == KIND_1D else nx * ny * batch


def flops(req):
    """5 N log2 N per transform; conv counts its forward and inverse transforms of the
    padded image (the filter spectra are resident)."""
    from .conv import good_size

    _, kind, n, nx, ny, batch, fid = req
    if kind == KIND_1D:
        return 5.0 * n * np.log2(n) * batch
    if kind == KIND_2D:
        return 5.0 * nx * ny * np.log2(nx * ny) * batch
    p = good_size(nx + 62)
    return 2 * 5.0 * p * p * np.log2(p * p)


class Arena:
    """Shared read-only inputs and a ring of outputs on the device for trace replay.

    Every request of a kind reads the same input region (uniform[-1,1] complex for FFTs,
    uniform[0,1] images for conv) and writes its own slice of an output ring, assigned at
    submission. While less than the ring's size of output is outstanding no two requests
    in flight share memory; a saturation replay (everything submitted at once) wraps the
    ring, so late requests overwrite early outputs -- timing is unaffected, and results
    are checked by the tests with buffers of their own.
    """

    def __init__(self, server, trace, ring_bytes, device="cuda"):
        import torch

        self.server = server
        dt = server.torch_dtype
        esz = torch.tensor([], dtype=dt).element_size()
        max_fft = max([2 * points(r) for r in trace if r[1] != KIND_CONV] + [2])
        g = torch.Generator(device="cpu").manual_seed(7)
        self.fft_in = (torch.rand(max_fft, generator=g) * 2 - 1).to(dt).to(device)
        side = max(SIDES_CONV)
        self.img_in = torch.rand(side * side, generator=g).to(device)
        self.ring = torch.empty(ring_bytes // esz, dtype=dt, device=device)
        self.esz = esz
        self.off = 0
        self.ring_bytes = ring_bytes

    def out_ptr(self, nbytes):
        nbytes = (nbytes + 255) // 256 * 256
        if self.off + nbytes > self.ring_bytes:
            self.off = 0
        p = self.ring.data_ptr() + self.off
        self.off += nbytes
        return p

    def pointers(self, req):
        _, kind, n, nx, ny, batch, _ = req
        if kind == KIND_CONV:
            return self.img_in.data_ptr(), self.out_ptr(nx * ny * 4)
        return self.fft_in.data_ptr(), self.out_ptr(2 * points(req) * self.esz)


def replay(server, trace, arena, time_scale=1.0, timeout_s=600.0):
    """Submit the trace at its arrival times (divided by time_scale), poll completions, and
    return per-request latency (ms) and the throughput over the run.

    Python's cyclic garbage collector is paused for the replay: a full collection over a
    large heap stops this loop for milliseconds, and the arrivals of that gap would then be
    submitted as one burst the trace does not contain. submit_lag_* report how late the
    submissions were against the trace's arrival times."""
    n = len(trace)
    done = {}
    lag = np.zeros(n)
    gc_on = gc.isenabled()
    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter()
        i = 0
        while i < n or len(done) < n:
            now = time.perf_counter() - t0
            while i < n and trace[i][0] / time_scale <= now:
                r = trace[i]
                ip, op = arena.pointers(r)
                server.submit(r[1], ip, op, tag=i, n=r[2], nx=r[3], ny=r[4], batch=r[5], filter_id=r[6])
                lag[i] = time.perf_counter() - t0 - r[0] / time_scale
                i += 1
            for c in server.poll():
                done[c.tag] = c
            if now > timeout_s:
                raise ServerError(f"trace replay timed out with {n - len(done)} requests outstanding")
            if i < n:
                nxt = trace[i][0] / time_scale - (time.perf_counter() - t0)
                if nxt > 2e-4:
                    time.sleep(min(nxt, 1e-3) - 1e-4)
            else:
                time.sleep(2e-4)
    finally:
        if gc_on:
            gc.enable()
    bad = [c for c in done.values() if c.status != 0]
    if bad:
        raise ServerError(f"{len(bad)} requests failed (first: tag {bad[0].tag} status {bad[0].status})")
    lat = np.array([(done[k].end_us - done[k].submit_us) / 1e3 for k in range(n)])
    first = min(c.submit_us for c in done.values())
    last = max(c.end_us for c in done.values())
    span = (last - first) / 1e6
    return {
        "requests": n,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "span_s": span,
        "req_per_s": n / span,
        "gflops": sum(flops(r) for r in trace) / span / 1e9,
        "submit_lag_p99_ms": float(np.percentile(lag, 99)) * 1e3,
        "submit_lag_max_ms": float(lag.max()) * 1e3,
    }


def register_filters(server, filters):
    for i, f in enumerate(filters):
        server.add_filter(i, f)


def warm(server, trace, arena):
    """Build every plan and conv operator the trace needs (one request per shape), outside
    any timed region."""
    seen, warm_reqs = set(), []
    for r in trace:
        key = (r[1], r[2], r[3], r[4], r[6] % 2 if r[1] == KIND_CONV else 0)
        if key not in seen:
            seen.add(key)
            warm_reqs.append(r)
    for k, r in enumerate(warm_reqs):
        ip, op = arena.pointers(r)
        server.submit(r[1], ip, op, tag=k, n=r[2], nx=r[3], ny=r[4], batch=r[5], filter_id=r[6])
        server.wait_idle(300.0)
    server.poll()
    return len(warm_reqs)
