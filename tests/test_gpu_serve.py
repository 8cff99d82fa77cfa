"""GPU tests of the continuous-batching request server (mfft_server_* in include/mfft.h).

  * every result of a mixed batch of 1D FFT / 2D FFT / conv2d requests, all submitted at
    once (so they overlap on the device, the small 1D ones share one work-queue launch
    and conv requests on one operator share a complex transform in pairs), equals the
    same request's result when it is run alone: FFT results bit for bit (the same kernels
    on the same data; tolerance 0), conv results within rel-L2 2e-5 (a paired image
    shares its transform's rounding with its partner) -- and each is within tolerance of
    the FP64 oracle (split fp16: rel-L2 <= 1e-3 for FFTs, <= 1e-4 for conv);
  * admission by HBM budget: a request that can never fit is rejected with
    MFFT_ERR_NOMEM, and the footprint never exceeds the budget while the rest run;
  * plain mode (fp32, one request at a time) against the oracle (rel-L2 <= 1e-5).
"""
import numpy as np
import pytest

import oracle
from tests.helpers import complex_input, from_device, rel_errors, to_device

MFFT_ERR_NOMEM = -4


def _filters():
    f = []
    for k, s in enumerate((3, 5, 9, 4)):
        f.append(oracle.uniform(s * s, 300 + k, 0.0, 1.0).reshape(s, s).astype(np.float32))
    return f


def _requests():
    """(kind, n, nx, ny, batch, filter_id) mixing coalesced small 1D, large 1D, 2D and conv."""
    from fft_demo_amd import serve

    K1, K2, KC = serve.KIND_1D, serve.KIND_2D, serve.KIND_CONV
    return [
        (K1, 256, 0, 0, 3, 0), (K1, 4096, 0, 0, 2, 0), (KC, 0, 128, 128, 1, 0), (K1, 1 << 16, 0, 0, 2, 0),
        (K2, 0, 256, 256, 2, 0), (K1, 512, 0, 0, 7, 0), (KC, 0, 96, 160, 1, 3), (K1, 1 << 14, 0, 0, 5, 0),
        (K1, 1024, 0, 0, 1, 0), (K2, 0, 512, 256, 1, 0), (KC, 0, 128, 128, 1, 2), (K1, 2048, 0, 0, 4, 0),
        (K1, 1 << 18, 0, 0, 1, 0), (KC, 0, 96, 160, 1, 1), (K1, 256, 0, 0, 64, 0), (K2, 0, 1024, 1024, 1, 0),
    ]


def _make_io(dev, server, reqs):
    import torch

    ins, outs, refs = [], [], []
    for i, (kind, n, nx, ny, b, fid) in enumerate(reqs):
        if kind == 3:
            img = oracle.uniform(nx * ny, 500 + i, 0.0, 1.0).reshape(nx, ny).astype(np.float32)
            ins.append(torch.from_numpy(img).to(dev))
            outs.append(torch.empty(nx, ny, dtype=torch.float32, device=dev))
            refs.append(img)
        else:
            shape = (b, n) if kind == 1 else (b, nx, ny)
            x = complex_input(shape, 500 + i)
            ins.append(to_device(x, server.torch_dtype, dev))
            outs.append(torch.empty_like(ins[-1]))
            refs.append(x)
    return ins, outs, refs


def _submit(server, reqs, ins, outs, idx):
    for i in idx:
        kind, n, nx, ny, b, fid = reqs[i]
        server.submit(kind, ins[i].data_ptr(), outs[i].data_ptr(), tag=i, n=n, nx=nx, ny=ny, batch=b, filter_id=fid)


def _collect(server, count):
    server.wait_idle(120.0)
    done = {c.tag: c for c in server.poll()}
    assert len(done) == count
    for c in done.values():
        assert c.status == 0, f"tag {c.tag} status {c.status}"
    return done


@pytest.mark.gpu
def test_server_mix_equals_alone(dev):
    import torch

    from fft_demo_amd import serve

    reqs = _requests()
    filt = _filters()
    with serve.Server(precision="split", streams=4, max_batch_1d=64, max_batch_2d=4, slab_bytes=256 << 20) as S:
        for k, f in enumerate(filt):
            S.add_filter(k, f)
        ins, outs, refs = _make_io(dev, S, reqs)
        print("mix: submitting", len(reqs), "requests", flush=True)
        _submit(S, reqs, ins, outs, range(len(reqs)))
        done = _collect(S, len(reqs))
        torch.cuda.synchronize()
        mixed = [o.clone() for o in outs]
        st = S.stats()
        print("stats", st)
        assert st["coalesced"] >= 2  # small 1D requests shared a work-queue launch
        assert st["paired"] >= 2  # two conv requests on one operator shared a transform
        for i in range(len(reqs)):
            outs[i].zero_()
            print("alone:", i, reqs[i], flush=True)
            _submit(S, reqs, ins, outs, [i])
            _collect(S, 1)
            torch.cuda.synchronize()
            if reqs[i][0] == 3:
                a, m = outs[i].double(), mixed[i].double()
                d = float((a - m).norm() / m.norm())
                print(f"conv alone vs mix rel_l2={d:.2e}")
                assert d <= 2e-5, f"request {i} {reqs[i]}"
            else:
                assert torch.equal(outs[i], mixed[i]), f"request {i} {reqs[i]} differs alone vs in the mix"
        for i, (kind, n, nx, ny, b, fid) in enumerate(reqs):
            if kind == 3:
                got = mixed[i].double().cpu().numpy()
                ref = oracle.conv2d_same(refs[i], filt[fid])
                l2 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
                print(f"conv {nx}x{ny} f{fid} rel_l2={l2:.2e}")
                assert l2 <= 1e-4
            else:
                y = from_device(mixed[i]) / done[i].out_scale
                ref = oracle.fft(refs[i]) if kind == 1 else oracle.fft2d(refs[i])
                l2, mx = rel_errors(y, ref)
                print(f"kind {kind} n={n} {nx}x{ny} b={b} rel_l2={l2:.2e} max={mx:.2e}")
                assert l2 <= 1e-3


@pytest.mark.gpu
def test_server_hbm_admission(dev):
    import torch

    from fft_demo_amd import serve

    slab = 32 << 20
    budget = 2 * slab + (24 << 20)
    with serve.Server(precision="split", streams=2, hbm_budget=budget, slab_bytes=slab, max_batch_1d=64) as S:
        n, b = 1 << 16, 16  # 2 * 4 MiB of I/O per request: at most 3 fit at once
        x = complex_input((b, n), 77)
        inp = to_device(x, torch.float16, dev)
        outs = [torch.empty_like(inp) for _ in range(12)]
        for i in range(12):
            S.submit(1, inp.data_ptr(), outs[i].data_ptr(), tag=i, n=n, batch=b)
        big = torch.empty(2 * (1 << 20) * 8, dtype=torch.float16, device=dev)  # 32 MiB in + out > budget
        S.submit(1, big.data_ptr(), big.data_ptr(), tag=99, n=1 << 20, batch=8)
        S.wait_idle(120.0)
        done = {c.tag: c for c in S.poll()}
        st = S.stats()
        print("stats", st, "budget", budget)
        assert done[99].status == MFFT_ERR_NOMEM
        assert all(done[i].status == 0 for i in range(12))
        assert st["peak_bytes"] <= budget
        assert st["inflight_bytes"] == 0
        torch.cuda.synchronize()
        ref = oracle.fft(x)
        for i in range(12):
            l2, _ = rel_errors(from_device(outs[i]) / done[i].out_scale, ref)
            assert l2 <= 1e-3


@pytest.mark.gpu
def test_server_plain_mode_fp32(dev):
    import torch

    from fft_demo_amd import serve

    reqs = [r for r in _requests() if r[0] != 3 or r[5] == 0][:8]
    filt = _filters()
    with serve.Server(plain=True, max_batch_1d=64, max_batch_2d=4, slab_bytes=256 << 20) as S:
        for k, f in enumerate(filt):
            S.add_filter(k, f)
        ins, outs, refs = _make_io(dev, S, reqs)
        _submit(S, reqs, ins, outs, range(len(reqs)))
        done = _collect(S, len(reqs))
        torch.cuda.synchronize()
        for i, (kind, n, nx, ny, b, fid) in enumerate(reqs):
            if kind == 3:
                ref = oracle.conv2d_same(refs[i], filt[fid])
                got = outs[i].double().cpu().numpy()
                l2 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            else:
                y = from_device(outs[i]) / done[i].out_scale
                ref = oracle.fft(refs[i]) if kind == 1 else oracle.fft2d(refs[i])
                l2, _ = rel_errors(y, ref)
            print(f"plain kind {kind} n={n} {nx}x{ny} b={b} rel_l2={l2:.2e}")
            assert l2 <= 1e-5


@pytest.mark.gpu
def test_server_rejects_unsupported_lengths_at_submit(dev):
    """A 1D request of length 7 (a prime factor outside {2, 3, 5, 17}) or 1 is refused at
    submission, so it can never join a coalesced work-queue launch; the valid small
    requests submitted around it all complete against the oracle."""
    import torch

    from fft_demo_amd import serve

    with serve.Server(precision="split", max_batch_1d=64) as S:
        xs, outs = [], []
        bad = torch.zeros(2 * 7 * 4, dtype=torch.float16, device=dev)
        for i, n in enumerate((256, 1024, 48, 4096)):
            x = complex_input((3, n), 700 + i)
            xs.append(x)
            xd = to_device(x, torch.float16, dev)
            outs.append((xd, torch.empty_like(xd)))
            S.submit(serve.KIND_1D, xd.data_ptr(), outs[-1][1].data_ptr(), tag=i, n=n, batch=3)
            for badn in (7, 1):
                with pytest.raises(serve.ServerError, match="length"):
                    S.submit(serve.KIND_1D, bad.data_ptr(), bad.data_ptr(), tag=100 + i, n=badn, batch=4)
        with pytest.raises(serve.ServerError, match="shape"):
            S.submit(serve.KIND_2D, bad.data_ptr(), bad.data_ptr(), tag=200, nx=7, ny=4, batch=1)
        with pytest.raises(serve.ServerError, match="filter"):
            S.submit(serve.KIND_CONV, bad.data_ptr(), bad.data_ptr(), tag=201, nx=8, ny=8, filter_id=5)
        S.wait_idle(60.0)
        done = {c.tag: c for c in S.poll()}
        assert sorted(done) == [0, 1, 2, 3] and all(c.status == 0 for c in done.values())
        torch.cuda.synchronize()
        for i, x in enumerate(xs):
            l2, mx = rel_errors(from_device(outs[i][1]) / done[i].out_scale, oracle.fft(x))
            assert l2 <= 1e-3 and mx <= 4e-3, (i, l2, mx)


@pytest.mark.gpu
def test_server_filter_added_after_conv_requests(dev):
    """A filter registered after conv requests have built the operator of its size-parity
    class (and a filter id registered again with new weights) is used by the next requests:
    the operator is rebuilt with the class's current filters. Results against the oracle."""
    import torch

    from fft_demo_amd import serve

    h, w = 96, 128
    img = oracle.uniform(h * w, 810, 0.0, 1.0).reshape(h, w).astype(np.float32)
    f3 = oracle.uniform(9, 811, 0.0, 1.0).reshape(3, 3).astype(np.float32)
    f5 = oracle.uniform(25, 812, 0.0, 1.0).reshape(5, 5).astype(np.float32)
    f3b = oracle.uniform(9, 813, -1.0, 1.0).reshape(3, 3).astype(np.float32)
    with serve.Server(precision="split") as S:
        S.add_filter(0, f3)
        xd = torch.from_numpy(img).to(dev)
        out = torch.empty_like(xd)

        def run(fid, filt, tag):
            S.submit(serve.KIND_CONV, xd.data_ptr(), out.data_ptr(), tag=tag, nx=h, ny=w, filter_id=fid)
            S.wait_idle(60.0)
            (c,) = S.poll()
            assert c.status == 0, c
            torch.cuda.synchronize()
            l2, mx = oracle.err_l2max(out.double().cpu().numpy(), oracle.conv2d_same(img, filt))
            print(f"filter {fid} {filt.shape}: rel_l2={l2:.2e} rel_max={mx:.2e}")
            assert l2 <= 1e-4 and mx <= 1e-3

        run(0, f3, 1)
        S.add_filter(1, f5)  # same (odd, odd) class, larger frame
        run(1, f5, 2)
        run(0, f3, 3)
        S.add_filter(0, f3b)  # new weights for an existing id
        run(0, f3b, 4)
        run(1, f5, 5)
