"""Limits and argument errors of the plan/exec entry points on the GPU.

The longest 1D lines of the reference's size list, 2^24 and 2^27 points (tcfftCreate's
largest cases, tcfft_half.cu:1296-1326), are too long for the FP64 oracle in a test, so they
are checked through properties that hold at any size: a forward then inverse round trip
returns the input, Parseval's identity, and an impulse's closed-form transform on sampled
frequencies (split fp16 operands, fp16 input and intermediates, fp32 output: at 2^27 points the
per-stage 1/16 scaling leaves outputs near 5e-5, where fp16 words would be subnormal). The
impulse has amplitude 2^14: the scaling divides a unit impulse by the length done so far, so
after two 256-point passes its 16-bit intermediate would be 2^-16, below fp16's normal range
(6.1e-5), where the words keep a fixed absolute precision (2^-24) instead of 11 bits.

Invalid sizes, batches and buffers are refused with an error code and a message (the
reference's switch keeps stale state for a size it has no case for, tcfft_half.cu:1250-1330).
"""
import ctypes

import numpy as np
import pytest

ERR_INVALID, ERR_UNSUPPORTED = -1, -2  # MFFT_ERR_* in include/mfft.h


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1 << 24, 1 << 27])
def test_longest_lines_round_trip_parseval_impulse(dev, n):
    import torch

    import fft_demo_amd as F

    fwd = F.Plan((n,), 1, precision="split_fp16", in_fmt="c16", mid_fmt="c16", out_fmt="c32")
    inv = F.Plan((n,), 1, precision="split_fp16", in_fmt="c32", mid_fmt="c16", out_fmt="c32", scale=0, direction=1)
    s = fwd.out_scale
    assert abs(s * n - 1.0) < 1e-12  # per-stage 1/radix: forward = DFT / n
    print(f"n = 2^{n.bit_length() - 1}: passes {[(p['kernel'], p['R']) for p in fwd.info['passes']]}")

    g = torch.Generator(device=dev).manual_seed(1234)
    x = (torch.rand((1, n, 2), generator=g, device=dev) * 2 - 1).half()
    X = fwd.exec(x)
    back = inv.exec(X)
    torch.cuda.synchronize()
    xf = x.float()
    rt = (torch.linalg.vector_norm(back - xf) / torch.linalg.vector_norm(xf)).item()
    # Parseval with the 1/n scaling: ||X||^2 = ||x||^2 / n
    par = (torch.sum(X.double() ** 2) * n / torch.sum(xf.double() ** 2)).item()
    print(f"round trip rel-L2 {rt:.3e}; Parseval ratio {par:.6f}")
    assert rt <= 2e-3
    assert abs(par - 1.0) <= 2e-3

    # impulse of amplitude A at n0: X[k] = A exp(-2 pi i n0 k / n) / n, on 65536 sampled
    # frequencies
    n0, amp = 12345 % n, 2.0 ** 14
    imp = torch.zeros((1, n, 2), dtype=torch.float16, device=dev)
    imp[0, n0, 0] = amp
    Y = fwd.exec(imp)
    torch.cuda.synchronize()
    k = np.random.default_rng(7).choice(n, 65536, replace=False)
    k[:3] = [0, 1, n - 1]
    kt = torch.from_numpy(k).to(dev)
    got = Y[0, kt].double().cpu().numpy()
    got = (got[:, 0] + 1j * got[:, 1]) * (n / amp)
    ref = np.exp(-2j * np.pi * ((n0 * k) % n) / n)
    err = np.abs(got - ref).max()
    print(f"impulse at {n0}: max |n X[k] / A - W^(n0 k)| over 65536 frequencies = {err:.3e}")
    assert err <= 5e-3
    for p in (fwd, inv):
        p.close()


@pytest.mark.gpu
def test_invalid_sizes_batches_and_buffers_are_refused(dev):
    import torch

    import fft_demo_amd as F
    from fft_demo_amd import _lib

    lib = _lib.load()
    d = _lib.Desc()
    assert lib.mfft_desc_default(ctypes.byref(d)) == 0
    h = ctypes.c_void_p()
    # sizes with a prime factor the planner has no stage for, degenerate sizes, empty batches
    for n, b, code in [(7 * 64, 4, ERR_UNSUPPORTED), (1, 4, None), (0, 4, ERR_INVALID), (256, 0, ERR_INVALID),
                       (256, -1, ERR_INVALID)]:
        rc = lib.mfft_plan_create_1d(ctypes.byref(h), n, b, ctypes.byref(d))
        assert rc != 0 and not h.value, (n, b, rc)
        if code is not None:
            assert rc == code, (n, b, rc)
        assert lib.mfft_last_error(), (n, b)
    rc = lib.mfft_plan_create_2d(ctypes.byref(h), 7 * 32, 64, 2, ctypes.byref(d))
    assert rc == ERR_UNSUPPORTED and not h.value
    # exec on a valid plan: partial batches outside 1..batch and null buffers
    p = F.Plan((256,), 8)
    x = p.empty_in(device=dev)
    y = p.empty_out(device=dev)
    s = torch.cuda.current_stream(dev).cuda_stream
    for b in (0, -3, 9):
        rc = lib.mfft_exec_ex(p._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), b, None,
                              ctypes.c_void_p(s))
        assert rc == ERR_INVALID, (b, rc)
        assert b"batch" in lib.mfft_last_error()
    assert lib.mfft_exec(p._h, None, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(s)) == ERR_INVALID
    assert lib.mfft_exec(p._h, ctypes.c_void_p(x.data_ptr()), None, ctypes.c_void_p(s)) == ERR_INVALID
    assert lib.mfft_exec(None, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                         ctypes.c_void_p(s)) == ERR_INVALID
    # the plan still works after the refused calls
    x.copy_(torch.ones_like(x))
    p.exec(x, y)
    torch.cuda.synchronize()
    assert torch.isfinite(y.float()).all()
    p.close()


@pytest.mark.gpu
def test_batch_chunks_are_equal(dev):
    """A workspace limit that holds 7 of 8 transforms runs the batch as two chunks of 4 (one
    chunk of 7 and one of 1 would leave most of the device idle on the second); the results
    equal those of a plan whose workspace holds the whole batch, bit for bit (the same kernels
    per transform)."""
    import torch

    import fft_demo_amd as F
    from tests.helpers import complex_input, to_device

    n, b = 65536, 8
    per = n * 4  # one transform's 16-bit intermediate
    small = F.Plan((n,), b, precision="split_fp16", mid_fmt="c16", split_hint=256, workspace_limit=7 * per)
    whole = F.Plan((n,), b, precision="split_fp16", mid_fmt="c16", split_hint=256, workspace_limit=16 * per)
    assert small.info["batch_chunk"] == 4 and whole.info["batch_chunk"] == 8
    assert small.info["launches_per_exec"] == 2 * len(small.info["passes"])
    x = to_device(complex_input((b, n), 31), small.in_dtype(), dev)
    y1, y2 = small.exec(x), whole.exec(x)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    small.close()
    whole.close()


@pytest.mark.gpu
def test_streaming_pass_beyond_32bit_offsets(dev):
    """The streaming kernel addresses through a buffer resource with 32-bit offsets when every
    offset of its launch fits, and with 64-bit arithmetic otherwise. 2176-point lines (a
    radix-17 streaming pass, then 128-point matrix-core passes) x 2^19 in one chunk reach
    4.56 GB of fp16 words (above 2^32 bytes); the same plan in 128 MiB chunks stays below.
    Both are checked on sampled lines against the oracle and against each other."""
    import torch

    import fft_demo_amd as F
    import oracle
    from tests.helpers import from_device, rel_errors

    n, b = 2176, 1 << 19
    one = F.Plan((n,), b, precision="split_fp16", mid_fmt="c16", workspace_limit=8 << 30)
    chunked = F.Plan((n,), b, precision="split_fp16", mid_fmt="c16")
    print(f"passes {[(p['kernel'], p['R']) for p in one.info['passes']]}, chunks {one.info['batch_chunk']} / "
          f"{chunked.info['batch_chunk']}", flush=True)
    assert one.info["passes"][0]["kernel"] == "stream_small" and one.info["passes"][0]["R"] == 17
    assert one.info["batch_chunk"] == b and chunked.info["batch_chunk"] * n * 4 <= 128 << 20
    g = torch.Generator(device=dev).manual_seed(2176)
    x = (torch.rand((b, n, 2), generator=g, device=dev) * 2 - 1).half()
    y1, y2 = one.exec(x), chunked.exec(x)
    torch.cuda.synchronize()
    lines = [0, 1, 12345, b // 2 + 7, b - 2, b - 1]
    xs = from_device(x[lines])
    ref = oracle.fft(xs, sign=-1) * one.out_scale
    for y in (y1, y2):
        l2, mx = rel_errors(from_device(y[lines]), ref)
        print(f"sampled lines vs oracle: rel_l2 {l2:.3e} rel_max {mx:.3e}", flush=True)
        assert l2 <= 1e-3 and mx <= 4e-3
    num = den = 0.0
    for s in range(0, b, 1 << 16):
        a1, a2 = y1[s:s + (1 << 16)].float(), y2[s:s + (1 << 16)].float()
        num += torch.sum((a1 - a2) ** 2).item()
        den += torch.sum(a2 ** 2).item()
    print(f"one chunk vs 128 MiB chunks: rel_l2 {(num / den) ** 0.5:.3e}", flush=True)
    assert (num / den) ** 0.5 <= 5e-4
    for p in (one, chunked):
        p.close()
