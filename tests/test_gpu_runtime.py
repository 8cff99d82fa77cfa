"""GPU tests of the runtime pieces: hipGraph-captured plans, the work-queue batcher
(many transforms of uneven length and batch count in one persistent launch) and the
exec variant with a caller-supplied workspace."""
import ctypes

import numpy as np
import pytest

import oracle
from tests.helpers import complex_input, from_device, rel_errors, to_device


@pytest.mark.gpu
def test_graph_replays_match_eager(dev):
    """Replays of a captured plan, with eager work and synchronisations in between, each
    equal the eager result bit for bit."""
    import torch

    import fft_demo_amd as F
    from fft_demo_amd import _lib as L

    lib = L.load()
    for shape, b in [((256,), 4096), ((65536,), 8), ((512, 512), 2)]:
        plan = F.Plan(shape, b, precision="split_fp16")
        x = to_device(complex_input((b, *shape), 3), torch.float16, dev)
        ref = plan.exec(x).clone()
        out = torch.empty_like(ref)
        s = torch.cuda.Stream(device=dev)
        g = ctypes.c_void_p()
        L.check(lib.mfft_graph_create(ctypes.byref(g), plan._h, ctypes.c_void_p(x.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s.cuda_stream)), "graph")
        for k in range(4):
            out.zero_()
            junk = torch.randn(1 << 20, device=dev).sum()  # eager work between replays
            torch.cuda.synchronize()
            L.check(lib.mfft_graph_launch(g, ctypes.c_void_p(s.cuda_stream)), "launch")
            s.synchronize()
            assert torch.equal(out, ref), (shape, k)
            del junk
        lib.mfft_graph_destroy(g)
        plan.close()


@pytest.mark.gpu
def test_batcher_uneven_lengths_one_launch(dev):
    import torch

    import fft_demo_amd as F
    from fft_demo_amd import _lib as L

    lib = L.load()
    desc = F.make_desc(precision="split_fp16")
    bh = ctypes.c_void_p()
    L.check(lib.mfft_batcher_create(ctypes.byref(bh), ctypes.byref(desc), 16), "batcher")
    cases = [(16, 3), (136, 5), (256, 100), (1000, 7), (2176, 2), (4096, 9), (32, 1), (512, 33)]
    xs, outs, items = [], [], (L.BatchItem * len(cases))()
    for i, (n, c) in enumerate(cases):
        x = complex_input((c, n), 40 + i)
        xd = to_device(x, torch.float16, dev)
        od = torch.empty_like(xd)
        xs.append((x, xd))
        outs.append(od)
        items[i].in_, items[i].out, items[i].n, items[i].count = xd.data_ptr(), od.data_ptr(), n, c
    s = torch.cuda.current_stream(dev)
    for rep in range(3):  # the job slots rotate
        L.check(lib.mfft_batcher_exec(bh, items, len(cases), ctypes.c_void_p(s.cuda_stream)), "exec")
    torch.cuda.synchronize()
    for (n, c), (x, _), od in zip(cases, xs, outs):
        y = from_device(od) * n  # per-stage 1/r scaling: out = DFT / n
        l2, mx = rel_errors(y, oracle.fft(x))
        print(f"batch item n={n} count={c}: rel_l2={l2:.2e} max={mx:.2e}")
        assert l2 <= 1e-3 and mx <= 4e-3, (n, c)
    lib.mfft_batcher_destroy(bh)


@pytest.mark.gpu
def test_exec_ex_partial_batch_own_workspace(dev):
    import torch

    import fft_demo_amd as F
    from fft_demo_amd import _lib as L

    lib = L.load()
    plan = F.Plan((65536,), 16, precision="split_fp16")
    ws = torch.empty(max(16, plan.info["workspace_bytes"]), dtype=torch.uint8, device=dev)
    x = complex_input((16, 65536), 5)
    xd = to_device(x, torch.float16, dev)
    out = torch.zeros_like(xd)
    s = torch.cuda.current_stream(dev)
    L.check(lib.mfft_exec_ex(plan._h, ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(out.data_ptr()), 5,
                             ctypes.c_void_p(ws.data_ptr()), ctypes.c_void_p(s.cuda_stream)), "exec_ex")
    torch.cuda.synchronize()
    y = from_device(out) / plan.out_scale
    l2, _ = rel_errors(y[:5], oracle.fft(x[:5]))
    assert l2 <= 1e-3
    assert float(out[5:].abs().max()) == 0.0  # transforms past the requested batch untouched


@pytest.mark.gpu
def test_two_streams_share_plan_workspace(dev):
    """One plan (two passes through its own workspace) called from two streams with no
    synchronisation in between, alternating several times: each exec is ordered after the
    previous one's use of the workspace, so every output equals the single-stream result
    bit for bit."""
    import torch

    import fft_demo_amd as F

    n, b = 65536, 8
    plan = F.Plan((n,), b, precision="split_fp16", mid_fmt="c16", split_hint=256)  # the two-pass plan
    assert plan.info["workspace_bytes"] > 0
    xs = [to_device(complex_input((b, n), 60 + k), torch.float16, dev) for k in range(4)]
    want = [plan.exec(x).clone() for x in xs]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)]
    outs = [torch.zeros_like(w) for w in want]
    for rep in range(3):
        for k, x in enumerate(xs):
            plan.exec(x, outs[k], stream=streams[k % 2])
        torch.cuda.synchronize()
        for k in range(4):
            assert torch.equal(outs[k], want[k]), (rep, k)
            outs[k].zero_()
    plan.close()


@pytest.mark.gpu
def test_two_streams_share_conv_operator(dev):
    """One conv operator (its spectra buffers) called from two streams back to back: results
    equal the single-stream ones bit for bit."""
    import torch

    from fft_demo_amd.conv import Conv2d

    h, w, k = 120, 96, 7
    op = Conv2d(2, 1, 1, h, w, k, k, mode="same")
    op.set_filters(torch.from_numpy(oracle.uniform(k * k, 8, 0.0, 1.0).reshape(1, 1, k, k)).to(dev, torch.float32))
    xs = [torch.from_numpy(oracle.uniform(2 * h * w, 90 + i, 0.0, 1.0).reshape(2, 1, h, w)).to(dev, torch.float32)
          for i in range(4)]
    want = [op.forward(x).clone() for x in xs]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)]
    outs = [torch.zeros_like(t) for t in want]
    for i, x in enumerate(xs):
        op.forward(x, outs[i], stream=streams[i % 2])
    torch.cuda.synchronize()
    for i in range(4):
        assert torch.equal(outs[i], want[i]), i
    op.close()
