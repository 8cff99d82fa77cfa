"""GPU parity of the batched 1D matrix-core FFT against the FP64 oracle.

Tolerances (relative L2 / relative max vs FP64, input already rounded to the boundary
format): split fp16 <= 1e-3 (the target accuracy bar), fp16 <= 4e-3, bf16 <= 3e-2,
plain fp32 <= 1e-5. out_scale (1 or 1/N) is divided out before comparing.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import complex_input, from_device, rel_errors, to_device

TOL = {"split_fp16": (1e-3, 4e-3), "fp16": (4e-3, 2e-2), "bf16": (3e-2, 1e-1), "fp32": (1e-5, 1e-4)}
FMT = {"split_fp16": "c16", "fp16": "c16", "bf16": "cbf16", "fp32": "c32"}


def _run(dev, n, batch, prec, direction=-1, seed=11, scale=None, **kw):
    import fft_demo_amd as F

    plan = F.Plan((n,), batch, precision=prec, direction=direction, scale=scale, **kw)
    x = complex_input((batch, n), seed, FMT[prec] if "in_fmt" not in kw else kw["in_fmt"])
    xd = to_device(x, plan.in_dtype(), dev)
    y = from_device(plan.exec(xd)) / plan.out_scale
    ref = oracle.fft(x, sign=direction)
    return y, ref, plan


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["split_fp16", "fp16", "bf16", "fp32"])
def test_n256_batch4096(dev, prec):
    """The reference's only working plan (tcfft_half.cu:1252-1255), N=256 x 4096."""
    y, ref, plan = _run(dev, 256, 4096, prec)
    l2, mx = rel_errors(y, ref)
    print(f"N=256 b=4096 {prec}: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["split_fp16", "fp16", "bf16", "fp32"])
@pytest.mark.parametrize("batch", [1, 13, 4093, 20000])
def test_n256_ragged_tiles(dev, prec, batch):
    """256-point row tiles hold 16 lines and order their stages per wave (each line stays in
    one wave); batches that leave a partial last tile, one tile, and more tiles than the
    persistent grid (several tiles per workgroup) against the oracle."""
    y, ref, plan = _run(dev, 256, batch, prec, seed=batch % 97)
    l2, mx = rel_errors(y, ref)
    print(f"N=256 b={batch} {prec}: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]


@pytest.mark.gpu
@pytest.mark.parametrize("n", [16, 32, 64, 128, 512, 1024, 2048, 4096, 8192, 65536])
def test_sizes_split(dev, n):
    batch = max(1, min(64, (1 << 20) // n))
    y, ref, plan = _run(dev, n, batch, "split_fp16")
    l2, mx = rel_errors(y, ref)
    print(f"N={n} b={batch} split: rel_l2={l2:.3e} rel_max={mx:.3e} passes={len(plan.info['passes'])}")
    assert l2 <= 1e-3 and mx <= 4e-3


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["split_fp16", "fp32"])
@pytest.mark.parametrize("n", [1 << 17, 1 << 18, 1 << 19, 1 << 20])
def test_long_lengths(dev, n, prec):
    """Lengths past one 65536-point four-step (the serving mix draws N up to 2^20)."""
    y, ref, plan = _run(dev, n, 3, prec, seed=n % 1000)
    l2, mx = rel_errors(y, ref)
    print(f"N={n} b=3 {prec}: rel_l2={l2:.3e} rel_max={mx:.3e} passes={len(plan.info['passes'])}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]


@pytest.mark.gpu
@pytest.mark.parametrize("n", [256, 4096, 65536])
def test_inverse_split(dev, n):
    y, ref, plan = _run(dev, n, 4, "split_fp16", direction=1)
    l2, mx = rel_errors(y, ref)
    assert l2 <= 1e-3 and mx <= 4e-3


@pytest.mark.gpu
def test_odd_batch_partial_tiles(dev):
    """Batch counts that leave a partial last tile."""
    for n, b in [(256, 37), (1024, 5), (64, 3)]:
        y, ref, _ = _run(dev, n, b, "split_fp16", seed=5)
        l2, _ = rel_errors(y, ref)
        assert l2 <= 1e-3, (n, b, l2)


@pytest.mark.gpu
def test_fp32_io_split(dev):
    """Split precision with fp32 input and output (no output quantisation)."""
    y, ref, _ = _run(dev, 4096, 8, "split_fp16", in_fmt="c32", out_fmt="c32")
    l2, mx = rel_errors(y, ref)
    print(f"split c32 io: {l2:.3e} {mx:.3e}")
    assert l2 <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["split_fp16", "fp32"])
@pytest.mark.parametrize("n", [48, 80, 96, 240, 576, 1088, 2176, 2560, 12288, 69632])
def test_mixed_radix(dev, n, prec):
    """odd * 2^a lengths: the odd factor (3, 5, 9, 15, 17) as a streaming pass, the power
    of two on the matrix-core kernels."""
    b = max(1, min(16, (1 << 18) // n))
    y, ref, plan = _run(dev, n, b, prec, seed=n % 89)
    l2, mx = rel_errors(y, ref)
    print(f"N={n} b={b} {prec}: rel_l2={l2:.3e} rel_max={mx:.3e} passes={[p['R'] for p in plan.info['passes']]}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]


@pytest.mark.gpu
@pytest.mark.parametrize("n", [240, 2176])
def test_mixed_radix_single_pass(dev, n):
    """split_hint = 1 keeps the single mixed-radix tile pass."""
    y, ref, plan = _run(dev, n, 8, "split_fp16", seed=3, split_hint=1)
    assert len(plan.info["passes"]) == 1
    l2, mx = rel_errors(y, ref)
    assert l2 <= 1e-3 and mx <= 4e-3


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["split_fp16", "fp16", "bf16", "fp32"])
@pytest.mark.parametrize("n", [64, 128, 8192, 1 << 14])
def test_short_line_kernels_all_precisions(dev, n, prec):
    """64- and 128-point passes (one radix-16 stage plus a radix-4/8 tail on the matrix-core
    kernels): single-pass 64/128 and the four-step 8192 = 128 x 64, 16384 = 128 x 128."""
    b = max(1, min(64, (1 << 18) // n))
    y, ref, plan = _run(dev, n, b, prec, seed=n % 97 + 3)
    l2, mx = rel_errors(y, ref)
    kinds = [p["kernel"] for p in plan.info["passes"]]
    print(f"N={n} b={b} {prec}: rel_l2={l2:.3e} rel_max={mx:.3e} passes={[p['R'] for p in plan.info['passes']]} {kinds}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]


@pytest.mark.gpu
@pytest.mark.parametrize("n,batch,mid,timed", [(65536, 8, "c16", True), (1 << 20, 2, None, False)])
def test_exec_pass_by_pass_equals_exec(dev, n, batch, mid, timed):
    """mfft_exec_pass run for every pass in order (the per-kernel timing entry point) leaves
    exactly the output of one mfft_exec (bit for bit), on a two-pass plan (65536 with the
    split asked for: the planner's default there is the one-pass on-chip kernel) and a longer
    plan; with start/stop events (kernel launched through hipExtLaunchKernel) each pass's
    window is a positive time."""
    import torch

    import fft_demo_amd as F
    from fft_demo_amd import _lib
    from fft_demo_amd.hiptime import EventPool

    plan = F.Plan((n,), batch, precision="split_fp16", mid_fmt=mid, split_hint=256 if n == 65536 else 0)
    x = to_device(complex_input((batch, n), 23, "c16"), plan.in_dtype(), dev)
    want = plan.exec(x)
    got = plan.empty_out(device=dev)
    got.zero_()
    s = torch.cuda.current_stream()
    lib = _lib.load()
    npass = len(plan.info["passes"])
    assert npass >= 2
    pool = EventPool(2 * npass)
    for k in range(npass):
        ev = (pool.arr[2 * k], pool.arr[2 * k + 1]) if timed else (None, None)
        assert lib.mfft_exec_pass(plan._h, k, x.data_ptr(), got.data_ptr(), s.cuda_stream, *ev) == 0
    assert lib.mfft_exec_pass(plan._h, npass, x.data_ptr(), got.data_ptr(), s.cuda_stream, None, None) == -1
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    if timed:
        ms = [pool.elapsed_ms(2 * k, 2 * k + 1) for k in range(npass)]
        print(f"pass windows (ms): {ms}")
        assert all(0 < t < 100 for t in ms)
    pool.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1 << 19, 1 << 20, 1 << 21])
@pytest.mark.parametrize("prec,mid", [("split_fp16", "c16"), ("fp16", None)])
def test_three_pass_lengths_narrow_intermediates(dev, n, prec, mid):
    """With 16-bit intermediates 2^20 and 2^21 plan three passes (first pass 2^(m-13), then
    8192 = 128 x 64) and 2^19 two (a 1024-point strided first pass, then 512 points);
    tolerance as for the precision."""
    y, ref, plan = _run(dev, n, 2, prec, mid_fmt=mid)
    l2, mx = rel_errors(y, ref)
    kinds = [(p["kernel"], p["R"]) for p in plan.info["passes"]]
    print(f"n=2^{n.bit_length() - 1} {prec} mid={mid}: rel_l2={l2:.2e} max={mx:.2e} passes={kinds}")
    if n == 1 << 19:
        assert kinds == [("fast_cols", 1024), ("fast_rows_tr", 512)], kinds
    else:
        assert len(kinds) == 3
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]


@pytest.mark.gpu
@pytest.mark.parametrize("prec,mid,l1", [("split_fp16", "c16", 1024), ("split_fp16", None, 512),
                                         ("fp16", None, 1024), ("fp32", None, 512)])
@pytest.mark.parametrize("n", [1 << 17, 1 << 18])
def test_default_long_first_pass(dev, n, prec, mid, l1):
    """2^17..2^19 plan a long strided first pass by default: 1024 points with 16-bit
    intermediates, 512 with 32-bit ones; the rest in one contiguous pass."""
    y, ref, plan = _run(dev, n, 3, prec, seed=n % 997, mid_fmt=mid)
    kinds = [(p["kernel"], p["R"]) for p in plan.info["passes"]]
    assert kinds == [("fast_cols", l1), ("fast_rows_tr", n // l1)], kinds
    l2, mx = rel_errors(y, ref)
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["split_fp16", "fp16", "fp32"])
@pytest.mark.parametrize("n,l1", [(1 << 19, 512), (1 << 20, 1024), (1 << 16, 1024)])
def test_four_step_long_strided_first_pass(dev, n, l1, prec):
    """Four-step splits whose strided first pass has 512 or 1024 points (the strided
    matrix-core kernel at 16 and 8 lines per tile, four-step twiddle fused in its store)."""
    y, ref, plan = _run(dev, n, 2, prec, seed=l1, split_hint=l1)
    kinds = [(p["kernel"], p["R"]) for p in plan.info["passes"]]
    assert kinds[0] == ("fast_cols", l1), kinds
    l2, mx = rel_errors(y, ref)
    print(f"N={n} split {l1} {prec}: rel_l2={l2:.3e} rel_max={mx:.3e} passes={kinds}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]
