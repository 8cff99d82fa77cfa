"""GPU parity of the on-chip kernel for 65536-point lines (fft_onchip: one workgroup per line,
the 256 x 256 four-step intermediate kept in LDS and registers), which the planner picks for
natural-order 65536-point lines with 16-bit words in, out and between the halves.

The planner takes the on-chip kernel for plans of at least 2/3 as many lines as the device
has compute units (171 on 256; fewer lines would leave it idle, and the two-pass plan is then
faster); such a plan keeps the two-pass plan for execs of fewer lines. Against the FP64 oracle
(the reference's check, main.cpp:164-198): split fp16, fp16 and bf16, forward and inverse,
per-stage and no scaling, 256 and 300 lines (more than the device runs at once), execs of 5
lines on such a plan (the companion two-pass plan), in place, digit-reversed input (tcfftExec's
order: gen_rev of {16, 16, 16, 16}); and against the two-pass plan of the same transform
(split_hint = 256: the strided 256-point pass with the four-step twiddle, then the contiguous
one), which computes the same decomposition in the same arithmetic with the intermediate in
HBM. Inputs: re/im uniform[-1,1] from SplitMix64, rounded to the input format.
Tolerances (relative L2 / relative max over all real components): split fp16 1e-3 / 4e-3,
fp16 4e-3 / 2e-2, bf16 3e-2 / 1e-1.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import complex_input, from_device, rel_errors, to_device

N = 65536
TOL = {"split_fp16": (1e-3, 4e-3), "fp16": (4e-3, 2e-2), "bf16": (3e-2, 1e-1)}
FMT = {"split_fp16": "c16", "fp16": "c16", "bf16": "cbf16"}


def _plan(prec, batch, **kw):
    import fft_demo_amd as F

    if prec == "split_fp16":
        kw.setdefault("mid_fmt", "c16")
    return F.Plan((N,), batch, precision=prec, **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["split_fp16", "fp16", "bf16"])
@pytest.mark.parametrize("direction", [-1, 1])
@pytest.mark.parametrize("batch", [256, 300])
def test_onchip_vs_oracle(dev, prec, direction, batch):
    plan = _plan(prec, batch, direction=direction)
    info = plan.info
    assert [(p["kernel"], p["R"]) for p in info["passes"]] == [("onchip", N)], info["passes"]
    assert info["launches_per_exec"] == 1
    x = complex_input((batch, N), 1000 + 7 * batch + direction, FMT[prec])
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
    l2, mx = rel_errors(y, oracle.fft(x, sign=direction))
    print(f"onchip {prec} dir={direction} b={batch}: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]
    plan.close()


@pytest.mark.gpu
def test_onchip_small_batches(dev):
    """A plan of few lines is the two-pass plan; an on-chip plan runs an exec of few lines on
    its companion two-pass plan (mfft_exec_ex with batch 5), against the oracle and bit for
    bit equal to the two-pass plan's own exec of the same lines."""
    import ctypes

    import torch

    from fft_demo_amd import _lib as L

    few = _plan("split_fp16", 5)
    assert [p["kernel"] for p in few.info["passes"]] == ["fast_cols", "fast_rows_tr"]
    big = _plan("split_fp16", 256)
    assert big.info["passes"][0]["kernel"] == "onchip"
    assert big.info["workspace_bytes"] >= few.info["workspace_bytes"] > 0  # the companion's
    x = complex_input((5, N), 71)
    xd = to_device(x, big.in_dtype(), dev)
    want = few.exec(xd)
    got = torch.zeros_like(want)
    s = torch.cuda.current_stream(dev)
    lib = L.load()
    L.check(lib.mfft_exec_ex(big._h, ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(got.data_ptr()), 5, None,
                             ctypes.c_void_p(s.cuda_stream)), "mfft_exec_ex")
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    l2, mx = rel_errors(from_device(got) / big.out_scale, oracle.fft(x))
    assert l2 <= 1e-3 and mx <= 4e-3
    few.close()
    big.close()


@pytest.mark.gpu
def test_onchip_unscaled(dev):
    """scale = 0 (no per-stage scaling): the output is the plain DFT."""
    plan = _plan("split_fp16", 256, scale=0)
    assert plan.info["passes"][0]["kernel"] == "onchip" and plan.out_scale == 1.0
    x = complex_input((256, N), 31, "c16", -1 / 64, 1 / 64)  # |DFT| <= N / 64 * sqrt(2): inside fp16
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev)))
    l2, mx = rel_errors(y, oracle.fft(x))
    print(f"onchip unscaled: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= 1e-3 and mx <= 4e-3
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["split_fp16", "bf16"])
def test_onchip_matches_two_pass(dev, prec):
    """The same transform through the two-pass plan (intermediate rounded to the same
    16-bit format in HBM): the two outputs agree to well within either's error bound."""
    import torch

    b = 256
    one = _plan(prec, b)
    two = _plan(prec, b, split_hint=256)
    assert [p["kernel"] for p in two.info["passes"]] == ["fast_cols", "fast_rows_tr"]
    x = complex_input((b, N), 55, FMT[prec])
    xd = to_device(x, one.in_dtype(), dev)
    y1 = from_device(one.exec(xd))
    y2 = from_device(two.exec(xd))
    torch.cuda.synchronize()
    l2, mx = rel_errors(y1, y2)
    ref = oracle.fft(x) * one.out_scale
    e1, _ = rel_errors(y1, ref)
    e2, _ = rel_errors(y2, ref)
    print(f"{prec}: onchip vs two-pass rel_l2={l2:.3e} rel_max={mx:.3e}; vs oracle {e1:.3e} / {e2:.3e}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]
    one.close()
    two.close()


@pytest.mark.gpu
def test_onchip_in_place(dev):
    """Input and output in one buffer (each workgroup reads its whole line before it writes
    any of it): equal bit for bit to the out-of-place result."""
    import torch

    b = 200
    plan = _plan("split_fp16", b)
    assert plan.info["passes"][0]["kernel"] == "onchip"
    x = complex_input((b, N), 61)
    xd = to_device(x, plan.in_dtype(), dev)
    want = plan.exec(xd.clone())
    plan.exec(xd, xd)
    torch.cuda.synchronize()
    assert torch.equal(xd, want)
    plan.close()


@pytest.mark.gpu
def test_onchip_impulses(dev):
    """Closed forms: a unit impulse at n0 transforms to exp(-2 pi i n0 k / N) (no scaling: every
    intermediate has magnitude 1); n0 chosen to exercise every digit position of the
    256 x 256 split."""
    plan = _plan("split_fp16", 256, scale=0)
    assert plan.info["passes"][0]["kernel"] == "onchip"
    n0s = [0, 1, 257, 40961]
    x = np.zeros((256, N), complex)
    for i, n0 in enumerate(n0s):
        x[i, n0] = 1.0
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev)))
    k = np.arange(N)
    for i, n0 in enumerate(n0s):
        ref = np.exp(-2j * np.pi * ((n0 * k) % N) / N)
        err = np.abs(y[i] - ref).max()
        print(f"impulse at {n0}: max abs error = {err:.3e}")
        assert err <= 2e-3
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,scale,batch", [("fp16", 0, 180), ("split_fp16", 0, 256), ("split_fp16", 1, 260)])
def test_onchip_digit_reversed(dev, prec, scale, batch):
    """Input stored in the digit-reversed order of four radix-16 digits (gen_rev of
    {16, 16, 16, 16}, the order tcfftExec reads at 65536 points): the on-chip kernel reads it
    with one 16-byte load per operand group; natural-order output against the oracle, and
    bit for bit equal to the natural-order plan on the same (unpermuted) line."""
    import torch

    plan = _plan(prec, batch, in_order=1, scale=scale)
    nat = _plan(prec, batch, scale=scale)
    assert [(p["kernel"], p["R"]) for p in plan.info["passes"]] == [("onchip", N)], plan.info["passes"]
    rev = oracle.gen_rev(N, [16, 16, 16, 16])
    assert np.array_equal(plan.input_permutation(), rev)
    x = complex_input((batch, N), 300 + batch, "c16", -1 / 8, 1 / 8)
    y = plan.exec(to_device(x[:, rev], plan.in_dtype(), dev))
    want = nat.exec(to_device(x, nat.in_dtype(), dev))
    torch.cuda.synchronize()
    assert torch.equal(y, want)
    l2, mx = rel_errors(from_device(y) / plan.out_scale, oracle.fft(x))
    print(f"onchip digit-reversed {prec} scale={scale} b={batch}: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]
    plan.close()
    nat.close()


# ---------------------------------------------------------------- 2048-point columns
@pytest.mark.gpu
@pytest.mark.parametrize("prec,nx,ny,batch,direction", [("split_fp16", 2048, 2048, 4, -1), ("split_fp16", 2048, 96, 64, 1),
                                                         ("fp16", 2048, 64, 96, -1), ("bf16", 2048, 32, 192, -1)])
def test_onchip_cols_2d_vs_oracle(dev, prec, nx, ny, batch, direction):
    """2D plans with 2048 rows and 16-bit words, with enough 32-column blocks to occupy 3/4
    of the compute units: the column axis is one pass of the on-chip column kernel (32
    adjacent columns per workgroup); against the oracle's 2D FFT."""
    import fft_demo_amd as F

    kw = dict(mid_fmt="c16") if prec == "split_fp16" else {}
    plan = F.Plan((nx, ny), batch, precision=prec, direction=direction, **kw)
    kinds = [p["kernel"] for p in plan.info["passes"]]
    assert kinds[-1] == "onchip_cols" and plan.info["passes"][-1]["R"] == 2048, kinds
    x = complex_input((batch, nx, ny), 900 + ny + batch, FMT[prec])
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
    l2, mx = rel_errors(y, oracle.fft2d(x, sign=direction))
    print(f"onchip_cols {prec} {nx}x{ny} b={batch} dir={direction}: passes {kinds} rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]
    plan.close()


@pytest.mark.gpu
def test_onchip_cols_matches_split_columns(dev):
    """The same 2048 x 2048 transform with the column axis split (split_hint = 128: a strided
    128-point matrix-core pass, then the streaming 16-point pass): the outputs agree well
    within either's error bound; a column count that is not a multiple of 32, and a plan of
    one image, keep the split."""
    import torch

    import fft_demo_amd as F

    b = 4
    one = F.Plan((2048, 2048), b, precision="split_fp16", mid_fmt="c16")
    assert [p["kernel"] for p in one.info["passes"]][-1] == "onchip_cols"
    two = F.Plan((2048, 2048), b, precision="split_fp16", mid_fmt="c16", split_hint=128)
    assert [p["kernel"] for p in two.info["passes"]][-1] != "onchip_cols"
    odd = F.Plan((2048, 48), 256, precision="split_fp16", mid_fmt="c16")
    assert "onchip_cols" not in [p["kernel"] for p in odd.info["passes"]]
    single = F.Plan((2048, 2048), 1, precision="split_fp16", mid_fmt="c16")
    assert "onchip_cols" not in [p["kernel"] for p in single.info["passes"]]
    single.close()
    x = complex_input((b, 2048, 2048), 77)
    xd = to_device(x, one.in_dtype(), dev)
    y1 = from_device(one.exec(xd))
    y2 = from_device(two.exec(xd))
    torch.cuda.synchronize()
    l2, mx = rel_errors(y1, y2)
    print(f"2048^2: one-pass columns vs split columns rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= 1e-3 and mx <= 4e-3
    for p in (one, two, odd):
        p.close()


@pytest.mark.gpu
def test_onchip_cols_impulses(dev):
    """Closed form per column: an impulse at row n0 of column c transforms (2D, unscaled) to
    exp(-2 pi i (n0 kx / 2048 + c ky / ny)); rows n0 chosen to exercise every digit of the
    16 x 128 split and columns on both halves of a 32-column block."""
    import fft_demo_amd as F

    nx, ny = 2048, 64
    plan = F.Plan((nx, ny), 96, precision="split_fp16", mid_fmt="c16", scale=0)
    assert plan.info["passes"][-1]["kernel"] == "onchip_cols"
    cases = [(0, 0), (1, 17), (129, 33), (1931, 62)]
    x = np.zeros((96, nx, ny), complex)
    for i, (n0, c0) in enumerate(cases):
        x[i, n0, c0] = 1.0
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev)))
    kx, ky = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    for i, (n0, c0) in enumerate(cases):
        ref = np.exp(-2j * np.pi * (((n0 * kx) % nx) / nx + ((c0 * ky) % ny) / ny))
        err = np.abs(y[i] - ref).max()
        print(f"impulse at ({n0}, {c0}): max abs error = {err:.3e}")
        assert err <= 2e-3
    plan.close()


# ---------------------------------------------------------------- 1D N = L1 x 2048
@pytest.mark.gpu
@pytest.mark.parametrize("prec,n,batch,direction", [("split_fp16", 1 << 20, 16, -1), ("split_fp16", 1 << 19, 32, 1),
                                                    ("split_fp16", 1 << 20, 13, 1), ("split_fp16", 1 << 19, 64, -1),
                                                    ("fp16", 1 << 20, 12, -1), ("bf16", 1 << 19, 24, -1)])
def test_onchip_rows_1d_vs_oracle(dev, prec, n, batch, direction):
    """1D lengths L1 x 2048 (2^19, 2^20) with 16-bit words and enough 32-row blocks to occupy
    3/4 of the compute units: pass A the strided L1-point matrix-core pass with the four-step
    twiddle, pass B the on-chip line kernel's rows kind (32 whole rows per workgroup, stored
    transposed); against the oracle."""
    import fft_demo_amd as F

    kw = dict(mid_fmt="c16") if prec == "split_fp16" else {}
    plan = F.Plan((n,), batch, precision=prec, direction=direction, **kw)
    kinds = [(p["kernel"], p["R"]) for p in plan.info["passes"]]
    assert kinds == [("fast_cols", n // 2048), ("onchip_cols", 2048)], kinds
    x = complex_input((batch, n), 1100 + n % 997 + batch, FMT[prec])
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
    l2, mx = rel_errors(y, oracle.fft(x, sign=direction))
    print(f"onchip rows {prec} n={n} b={batch} dir={direction}: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]
    plan.close()


@pytest.mark.gpu
def test_onchip_rows_matches_multipass_and_small_execs(dev):
    """2^20 x 16: the two-pass plan with the on-chip rows agrees with the three-pass plan
    (split_hint = 128) well within either's error bound; an exec of 3 lines on it runs its
    companion multi-pass plan (bit for bit equal to a plan of 3 lines)."""
    import ctypes

    import torch

    import fft_demo_amd as F
    from fft_demo_amd import _lib as L

    n, b = 1 << 20, 16
    one = F.Plan((n,), b, precision="split_fp16", mid_fmt="c16")
    three = F.Plan((n,), b, precision="split_fp16", mid_fmt="c16", split_hint=128)
    few = F.Plan((n,), 3, precision="split_fp16", mid_fmt="c16")
    assert [p["kernel"] for p in one.info["passes"]][-1] == "onchip_cols"
    assert "onchip_cols" not in [p["kernel"] for p in three.info["passes"] + few.info["passes"]]
    x = complex_input((b, n), 1201)
    xd = to_device(x, one.in_dtype(), dev)
    y1 = from_device(one.exec(xd))
    y3 = from_device(three.exec(xd))
    l2, mx = rel_errors(y1, y3)
    print(f"2^20: two passes (on-chip rows) vs three: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= 1e-3 and mx <= 4e-3
    want = few.exec(xd[:3].contiguous())
    got = torch.zeros_like(want)
    s = torch.cuda.current_stream(dev)
    lib = L.load()
    L.check(lib.mfft_exec_ex(one._h, ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(got.data_ptr()), 3, None,
                             ctypes.c_void_p(s.cuda_stream)), "mfft_exec_ex")
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    for p in (one, three, few):
        p.close()
