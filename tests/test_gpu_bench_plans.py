"""GPU parity of the exact plans bench.py reports, at the benchmark's own sizes and inputs,
against the FP64 oracle (the reference's accuracy check, main.cpp:164-198, on the
workloads the throughput numbers are quoted for).

  * 1D N = 65536 x 256, split fp16 with fp16 intermediates (the headline plan: the
    planner's one-pass on-chip kernel, and the two-pass schedule -- a strided 256-point pass
    with the four-step twiddle, then a contiguous 256-point pass with a transposed store --
    asked for with split_hint = 256), and under the autotuned schedule; the inverse; fp16
    and bf16 plans of the same size;
  * 2D 1024 x 1024 x 64 with 10% of the entries x100, split fp16 with fp16 intermediates:
    forward against the oracle, the inverse of that output against the oracle's inverse,
    and the round trip against the input;
  * FFT 'same' convolution 2048 x 2048 x 16, 31 x 31 Gaussian filter, pad 2176, fp16 image
    spectra (the bench's conv operator) on sampled row bands of two images;
  * the largest request shapes of the serving mix through the server (split fp16, fp16
    intermediates and conv spectra): 1D 2^20 x 64, 2D 2048^2 x 16, conv 2048^2 with the
    63 x 63 filter.

Inputs are the bench's: re/im uniform[-1,1] from SplitMix64 seed 42, rounded to fp16
(input quantisation is not counted as error). Tolerances (relative L2 / relative max over
all real components): split fp16 1e-3 / 4e-3, fp16 4e-3 / 2e-2, bf16 3e-2 / 1e-1; conv
with fp16 spectra 1e-3 / 1e-2 per sampled band.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import complex_input, from_device, rel_errors, to_device

TOL = {"split_fp16": (1e-3, 4e-3), "fp16": (4e-3, 2e-2), "bf16": (3e-2, 1e-1)}
FMT = {"split_fp16": "c16", "fp16": "c16", "bf16": "cbf16"}


def _kinds(plan):
    return [(p["kernel"], p["R"], p["in_elem_bytes"], p["out_elem_bytes"]) for p in plan.info["passes"]]


def _spiky_input(shape, batch, seed=42):
    """bench.py's 2D input: uniform[-1,1], 10% of the complex entries (mask from seed + 1)
    x100, rounded to fp16."""
    n = int(np.prod(shape))
    v = oracle.uniform(batch * n * 2, seed).reshape(batch * n, 2)
    mask = oracle.uniform(batch * n, seed + 1, 0.0, 1.0) < 0.1
    v[mask] *= 100.0
    v = oracle.round_fp16(v).reshape(batch, *shape, 2)
    return v[..., 0] + 1j * v[..., 1]


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["planner", "two_pass", "tuned"])
def test_c2_split_c16_mid_full_batch(dev, schedule):
    """The headline transform on the whole 65536 x 256 batch, every transform against the
    oracle: the planner's plan (one pass of the on-chip kernel, c16 -> c16), the two-pass
    plan (fast_cols 256 then fast_rows_tr 256, c16 throughout) and the autotuned one."""
    import fft_demo_amd as F
    from fft_demo_amd import tune

    n, batch = 65536, 256
    kw = dict(precision="split_fp16", mid_fmt="c16")
    if schedule == "tuned":
        e = tune.autotune((n,), batch, **kw)
        print("tuned schedule", e)
        kw["split_hint"] = tune.SPLIT_TUNED
    elif schedule == "two_pass":
        kw["split_hint"] = 256
    plan = F.Plan((n,), batch, **kw)
    kinds = _kinds(plan)
    print("passes", kinds, "chunk", plan.info["batch_chunk"])
    if schedule == "planner":
        assert kinds == [("onchip", n, 4, 4)], kinds
    elif schedule == "two_pass":
        assert kinds == [("fast_cols", 256, 4, 4), ("fast_rows_tr", 256, 4, 4)], kinds
    assert all(k[2] == 4 and k[3] == 4 for k in kinds)  # fp16 buffers throughout
    x = complex_input((batch, n), 42)
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
    l2, mx = rel_errors(y, oracle.fft(x))
    print(f"65536 x 256 split c16 mid {schedule}: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= 1e-3 and mx <= 4e-3
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("split", [0, 256])
def test_c2_split_c16_mid_inverse(dev, split):
    import fft_demo_amd as F

    n, batch = 65536, 256
    plan = F.Plan((n,), batch, precision="split_fp16", mid_fmt="c16", direction=1, split_hint=split)
    want = [("onchip", n)] if split == 0 else [("fast_cols", 256), ("fast_rows_tr", 256)]
    assert [k[:2] for k in _kinds(plan)] == want
    x = complex_input((batch, n), 43)
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
    l2, mx = rel_errors(y, oracle.fft(x, sign=1))
    print(f"65536 x {batch} split c16 mid inverse: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= 1e-3 and mx <= 4e-3
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_c2_half_precisions_full_batch(dev, prec):
    """The fp16 / bf16 comparators of the 65536 x 256 line (their default 16-bit
    intermediates, per-stage 1/16 scaling as the bench runs them)."""
    import fft_demo_amd as F

    n, batch = 65536, 256
    plan = F.Plan((n,), batch, precision=prec, scale=1)
    x = complex_input((batch, n), 42, FMT[prec])
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
    l2, mx = rel_errors(y, oracle.fft(x))
    print(f"65536 x 256 {prec}: rel_l2={l2:.3e} rel_max={mx:.3e} passes={_kinds(plan)}")
    assert l2 <= TOL[prec][0] and mx <= TOL[prec][1]
    plan.close()


@pytest.mark.gpu
def test_c3_split_c16_mid_round_trip_full_batch(dev):
    """2D 1024 x 1024 x 64 with 10% of entries x100: forward (1/radix per stage), inverse
    (unscaled) of the forward's output, each against the oracle on its own input, and the
    round trip against the input."""
    import torch

    import fft_demo_amd as F

    shape, batch = (1024, 1024), 64
    fwd = F.Plan(shape, batch, precision="split_fp16", mid_fmt="c16")
    inv = F.Plan(shape, batch, precision="split_fp16", mid_fmt="c16", scale=0, direction=1)
    print("forward passes", _kinds(fwd), "inverse passes", _kinds(inv))
    assert all(k[2] == 4 and k[3] == 4 for k in _kinds(fwd) + _kinds(inv))
    x = _spiky_input(shape, batch)
    xd = to_device(x, fwd.in_dtype(), dev)
    yd = fwd.exec(xd)
    del xd
    zd = inv.exec(yd)
    torch.cuda.synchronize()
    y = from_device(yd)
    del yd
    l2, mx = rel_errors(y / fwd.out_scale, oracle.fft2d(x))
    print(f"1024^2 x 64 forward: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= 1e-3 and mx <= 4e-3
    z = from_device(zd)
    del zd
    l2, mx = rel_errors(z / inv.out_scale, oracle.fft2d(y, sign=1))
    print(f"1024^2 x 64 inverse of the forward's output: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= 1e-3 and mx <= 4e-3
    l2, mx = rel_errors(z, x)  # forward scaled 1/N, inverse unscaled: z ~ x
    print(f"1024^2 x 64 round trip vs input: rel_l2={l2:.3e} rel_max={mx:.3e}")
    assert l2 <= 1e-3 and mx <= 4e-3
    fwd.close()
    inv.close()


def _gauss31():
    k = 31
    c = (k - 1) / 2
    yy, xx = np.mgrid[0:k, 0:k]
    g = np.exp(-((yy - c) ** 2 + (xx - c) ** 2) / (2 * 25.0))
    g /= g.sum()
    g *= 1 + oracle.uniform(k * k, 43, -0.01, 0.01).reshape(k, k)
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["spec_fp16", "mid_fp16"])
def test_c4_conv_same_fp16_spectra_pad2176(dev, mode):
    """The bench's conv operators: 16 images of 2048 x 2048 (U[0,1], seed 42), one 31 x 31
    Gaussian filter, pad 2176, split fp16 transforms with fp16 image spectra (spec_fp16) or
    fp32 image spectra and fp16 pass boundaries (mid_fp16). Row bands at the top, middle and
    bottom (all columns: both side borders) of images 0 and 15."""
    import torch

    from fft_demo_amd.conv import Conv2d

    B, H, W, K = 16, 2048, 2048, 31
    img = oracle.uniform(B * H * W, 42, 0.0, 1.0).astype(np.float32).reshape(B, 1, H, W)
    g = _gauss31()
    op = Conv2d(B, 1, 1, H, W, K, K, mode="same", precision="split_fp16", **{mode: True})
    assert tuple(op.pad) == (2176, 2176)
    op.set_filters(torch.from_numpy(g.astype(np.float32)).to(dev).reshape(1, 1, K, K).contiguous())
    out = op.forward(torch.from_numpy(img).to(dev))
    torch.cuda.synchronize()
    for b in (0, B - 1):
        for r0 in (0, 1016, 2032):
            ref = oracle.conv2d_same(img[b, 0].astype(np.float64), g, rows=(r0, r0 + 16))
            got = out[b, 0, r0:r0 + 16].double().cpu().numpy()
            l2, mx = oracle.err_l2max(got, ref)
            print(f"conv 2048^2 {mode} image {b} rows {r0}..{r0 + 15}: rel_l2={l2:.2e} rel_max={mx:.2e}")
            assert l2 <= 1e-3 and mx <= 1e-2
    op.close()


@pytest.mark.gpu
def test_serve_largest_mix_shapes(dev):
    """One request each of the serving mix's largest shapes, submitted together to the
    server the bench measures (split fp16, fp16 intermediates and conv spectra, 32 resident
    filters): 1D 2^20 x 64 and 2D 2048^2 x 16 on every transform, conv 2048^2 with the
    63 x 63 filter (id 30) on row bands; then each again alone, bit for bit equal (FFT) or
    within rel-L2 2e-5 (conv) of its result in the mix."""
    import torch

    from fft_demo_amd import serve

    filters = serve.make_filters()
    assert filters[30].shape == (63, 63)
    reqs = [(serve.KIND_1D, 1 << 20, 0, 0, 64, 0), (serve.KIND_2D, 0, 2048, 2048, 16, 0),
            (serve.KIND_CONV, 0, 2048, 2048, 1, 30)]
    with serve.Server(precision="split", mid_fmt="c16") as S:
        serve.register_filters(S, filters)
        ins, outs, refs = [], [], []
        for i, (kind, n, nx, ny, b, fid) in enumerate(reqs):
            if kind == serve.KIND_CONV:
                im = oracle.uniform(nx * ny, 900 + i, 0.0, 1.0).reshape(nx, ny).astype(np.float32)
                ins.append(torch.from_numpy(im).to(dev))
                outs.append(torch.empty(nx, ny, dtype=torch.float32, device=dev))
                refs.append(im)
            else:
                shape = (b, n) if kind == serve.KIND_1D else (b, nx, ny)
                x = complex_input(shape, 900 + i)
                ins.append(to_device(x, torch.float16, dev))
                outs.append(torch.empty_like(ins[-1]))
                refs.append(x)

        def submit(idx):
            for i in idx:
                kind, n, nx, ny, b, fid = reqs[i]
                S.submit(kind, ins[i].data_ptr(), outs[i].data_ptr(), tag=i, n=n, nx=nx, ny=ny, batch=b,
                         filter_id=fid)
            S.wait_idle(300.0)
            done = {c.tag: c for c in S.poll()}
            assert sorted(done) == sorted(idx) and all(c.status == 0 for c in done.values()), done
            torch.cuda.synchronize()
            return done

        print("submitting", reqs, flush=True)
        done = submit(range(len(reqs)))
        mixed = [o.clone() for o in outs]
        for i, (kind, n, nx, ny, b, fid) in enumerate(reqs):
            if kind == serve.KIND_CONV:
                for r0 in (0, 1016, 2032):
                    ref = oracle.conv2d_same(refs[i].astype(np.float64), filters[fid].astype(np.float64),
                                             rows=(r0, r0 + 16))
                    l2, mx = oracle.err_l2max(mixed[i][r0:r0 + 16].double().cpu().numpy(), ref)
                    print(f"serve conv 2048^2 63x63 rows {r0}..: rel_l2={l2:.2e} rel_max={mx:.2e}")
                    assert l2 <= 1e-3 and mx <= 1e-2
            else:
                y = from_device(mixed[i]) / done[i].out_scale
                ref = oracle.fft(refs[i]) if kind == serve.KIND_1D else oracle.fft2d(refs[i])
                l2, mx = rel_errors(y, ref)
                print(f"serve kind {kind} n={n} {nx}x{ny} b={b}: rel_l2={l2:.2e} rel_max={mx:.2e}")
                assert l2 <= 1e-3 and mx <= 4e-3
        for i in range(len(reqs)):
            outs[i].zero_()
            submit([i])
            if reqs[i][0] == serve.KIND_CONV:
                d = float((outs[i].double() - mixed[i].double()).norm() / mixed[i].double().norm())
                assert d <= 2e-5, d
            else:
                assert torch.equal(outs[i], mixed[i]), f"request {reqs[i]} differs alone vs in the mix"
