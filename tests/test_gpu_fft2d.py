"""GPU parity of the batched 2D matrix-core FFT ([batch][nx][ny], y contiguous) against
the FP64 oracle (rows then columns, main.cpp:82-94 convention).

Tolerances vs FP64 on the fp16-rounded input: split fp16 rel-L2 <= 1e-3 and rel-max
<= 4e-3; fp16 rel-L2 <= 4e-3.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import complex_input, from_device, rel_errors, to_device


def _plan(nx, ny, b, **kw):
    import fft_demo_amd as F

    return F.Plan((nx, ny), b, **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny,b", [(256, 256, 4), (64, 128, 3), (512, 512, 2), (1024, 1024, 2), (2048, 2048, 1),
                                     (16, 4096, 2), (4096, 32, 1), (2176, 2176, 1), (576, 640, 2), (96, 48, 3)])
def test_2d_split(dev, nx, ny, b):
    plan = _plan(nx, ny, b, precision="split_fp16")
    x = complex_input((b, nx, ny), nx + ny)
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
    l2, mx = rel_errors(y, oracle.fft2d(x))
    print(f"2D {nx}x{ny} b={b} split rel_l2={l2:.2e} max={mx:.2e} passes={[p['R'] for p in plan.info['passes']]}")
    assert l2 <= 1e-3 and mx <= 4e-3


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("split_fp16", 1e-3), ("fp16", 4e-3)])
def test_2d_1024_dynamic_range_round_trip(dev, prec, tol):
    """1024x1024, uniform[-1,1] with 10% of entries scaled x100; forward then inverse."""
    import torch

    b = 2
    x = complex_input((b, 1024, 1024), 77, "c32")
    mask = oracle.uniform(b * 1024 * 1024, 78, 0, 1).reshape(b, 1024, 1024) < 0.1
    x = np.where(mask, 100 * x, x)
    x = oracle.round_fp16(np.stack([x.real, x.imag], -1)).reshape(b, 1024, 1024, 2)
    x = x[..., 0] + 1j * x[..., 1]
    fwd = _plan(1024, 1024, b, precision=prec, scale=1)  # per-stage 1/radix: out = DFT/N
    inv = _plan(1024, 1024, b, precision=prec, scale=0, direction=1)
    xd = to_device(x, fwd.in_dtype(), dev)
    X = fwd.exec(xd)
    ref = oracle.fft2d(x)
    l2f, mxf = rel_errors(from_device(X) / fwd.out_scale, ref)
    back = inv.exec(X)
    torch.cuda.synchronize()
    l2r, mxr = rel_errors(from_device(back), x)
    print(f"1024^2 x100 {prec}: forward rel_l2={l2f:.2e} max={mxf:.2e}; round trip rel_l2={l2r:.2e} max={mxr:.2e}")
    assert l2f <= tol and l2r <= 2 * tol


@pytest.mark.gpu
def test_2d_inverse_and_fp32(dev):
    for prec, tol, fmt in [("split_fp16", 1e-3, "c16"), ("fp32", 1e-5, "c32")]:
        plan = _plan(256, 512, 2, precision=prec, direction=1)
        x = complex_input((2, 256, 512), 9, fmt)
        y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
        l2, _ = rel_errors(y, oracle.fft2d(x, sign=1))
        assert l2 <= tol, (prec, l2)
        plan = _plan(1088, 576, 1, precision=prec, direction=1)  # 17 * 64 by 9 * 64
        x = complex_input((1, 1088, 576), 10, fmt)
        y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
        l2, _ = rel_errors(y, oracle.fft2d(x, sign=1))
        assert l2 <= tol, (prec, l2)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp16", 4e-3), ("bf16", 3e-2), ("fp32", 1e-5)])
def test_2d_short_axes_other_precisions(dev, prec, tol):
    """2D with 64/128-point axes (rows on the contiguous kind, columns on the strided kind)."""
    fmt = {"fp16": "c16", "bf16": "cbf16", "fp32": "c32"}[prec]
    for nx, ny in [(64, 128), (128, 64), (2176, 128)]:
        plan = _plan(nx, ny, 2, precision=prec)
        x = complex_input((2, nx, ny), nx * 3 + ny, fmt)
        y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
        l2, _ = rel_errors(y, oracle.fft2d(x))
        print(f"2D {nx}x{ny} {prec}: rel_l2={l2:.2e} kernels={[p['kernel'] for p in plan.info['passes']]}")
        assert l2 <= tol, (nx, ny, prec, l2)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("split_fp16", 1e-3), ("fp16", 4e-3), ("bf16", 3e-2), ("fp32", 1e-5)])
def test_2d_long_columns_one_strided_pass(dev, prec, tol):
    """Column axes of 512 and 1024 points in ONE pass of the strided matrix-core kernel
    (16 and 8 lines per tile), full and ragged line counts (ny not a multiple of the tile's
    lines), forward and inverse."""
    fmt = {"split_fp16": "c16", "fp16": "c16", "bf16": "cbf16", "fp32": "c32"}[prec]
    for nx, ny, b, sign in [(1024, 40, 2, -1), (1024, 36, 1, 1), (512, 24, 2, -1), (512, 20, 1, 1),
                            (1024, 1024, 1, -1)]:
        plan = _plan(nx, ny, b, precision=prec, direction=sign)
        kinds = [(p["kernel"], p["R"]) for p in plan.info["passes"]]
        assert kinds[-1] == ("fast_cols", nx), kinds
        x = complex_input((b, nx, ny), nx + 7 * ny, fmt)
        y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
        l2, mx = rel_errors(y, oracle.fft2d(x, sign=sign))
        print(f"2D {nx}x{ny} b={b} {prec} sign={sign}: rel_l2={l2:.2e} max={mx:.2e} passes={kinds}")
        assert l2 <= tol, (nx, ny, prec, l2)


@pytest.mark.gpu
@pytest.mark.parametrize("mid", [None, "c16"])
def test_2d_long_columns_split_hint_keeps_split(dev, mid):
    """split_hint on a 1024-point column axis still splits it (64 on the strided kernel,
    16 streamed): the tuner's other candidate."""
    plan = _plan(1024, 64, 1, precision="split_fp16", split_hint=64, mid_fmt=mid)
    kinds = [(p["kernel"], p["R"]) for p in plan.info["passes"]]
    assert ("fast_cols", 64) in kinds and len(kinds) == 3, kinds
    x = complex_input((1, 1024, 64), 5)
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
    l2, mx = rel_errors(y, oracle.fft2d(x))
    assert l2 <= 1e-3 and mx <= 4e-3, (l2, mx)
