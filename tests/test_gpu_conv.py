"""GPU parity of the FFT-convolution operator against the direct FP64 convolution oracle
(oracle.conv2d_same: 'same' convolution; oracle.corr2d_valid: the reference's 'valid'
cross-correlation, main.cpp:97-116).

Tolerance: rel-L2 <= 1e-4 and rel-max <= 1e-3 of each output map (split-fp16 transforms,
fp32 spectra, fp32 images), measured on every pixel for small images and on sampled
rows (top, middle, bottom bands) for 2048 x 2048.
"""
import numpy as np
import pytest

import oracle


def _filters(F, C, k, seed):
    return oracle.uniform(F * C * k * k, seed, 0.0, 1.0).reshape(F, C, k, k)


def _conv(dev, B, C, F, h, w, fh, fw, mode, seed, pad=None, dtype="float32"):
    import torch

    from fft_demo_amd.conv import Conv2d

    img = oracle.uniform(B * C * h * w, seed, 0.0, 1.0).reshape(B, C, h, w)
    filt = oracle.uniform(F * C * fh * fw, seed + 1, -1.0, 1.0).reshape(F, C, fh, fw)
    td = getattr(torch, dtype)
    op = Conv2d(B, C, F, h, w, fh, fw, mode=mode, pad=pad)
    op.set_filters(torch.from_numpy(filt).to(device=dev, dtype=td).contiguous())
    out = op.forward(torch.from_numpy(img).to(device=dev, dtype=td).contiguous())
    torch.cuda.synchronize()
    return img, filt, out.double().cpu().numpy(), op


@pytest.mark.gpu
def test_conv_same_multichannel_odd_batch(dev):
    B, C, F, h, w, fh, fw = 3, 2, 2, 64, 80, 7, 5
    img, filt, out, op = _conv(dev, B, C, F, h, w, fh, fw, "same", 11)
    assert out.shape == (B, F, h, w)
    for b in range(B):
        for f in range(F):
            ref = sum(oracle.conv2d_same(img[b, c], filt[f, c]) for c in range(C))
            l2, mx = oracle.err_l2max(out[b, f], ref)
            assert l2 <= 1e-4 and mx <= 1e-3, (b, f, l2, mx)
    print("pad", op.pad)


@pytest.mark.gpu
def test_conv_valid_correlation_reference_config(dev):
    """main.cpp:320-352: 256x256 signal, 13x13 filters, 16 filters, 1 channel."""
    B, C, F, h, w, k = 1, 1, 16, 256, 256, 13
    img, filt, out, op = _conv(dev, B, C, F, h, w, k, k, "valid_corr", 21, dtype="float64")
    assert out.shape == (1, F, h - k + 1, w - k + 1)
    worst = 0.0
    for f in range(F):
        ref = oracle.corr2d_valid(img[0, 0], filt[f, 0])
        l2, mx = oracle.err_l2max(out[0, f], ref)
        worst = max(worst, l2)
        assert l2 <= 1e-4 and mx <= 1e-3, (f, l2, mx)
    print(f"valid corr 256^2 x 16 filters: worst rel_l2={worst:.2e}, pad={op.pad}")


@pytest.mark.gpu
def test_conv_same_2048_gaussian_pad2176(dev):
    """2048x2048 fp32 U[0,1] images, 31x31 Gaussian (sigma 5, sum 1) x (1 +- 1% noise),
    padding 2176 = 2048 + 30 rounded up."""
    import torch

    from fft_demo_amd.conv import Conv2d, good_size

    assert good_size(2078) == 2176
    B, k = 2, 31
    yy, xx = np.mgrid[:k, :k] - (k - 1) / 2
    g = np.exp(-(xx ** 2 + yy ** 2) / (2 * 5.0 ** 2))
    g /= g.sum()
    g *= 1 + oracle.uniform(k * k, 5, -0.01, 0.01).reshape(k, k)
    img = oracle.uniform(B * 2048 * 2048, 6, 0.0, 1.0).reshape(B, 1, 2048, 2048).astype(np.float32)
    op = Conv2d(B, 1, 1, 2048, 2048, k, k, mode="same")
    assert op.pad == (2176, 2176)
    op.set_filters(torch.from_numpy(g.reshape(1, 1, k, k)).to(dev, torch.float32).contiguous())
    out = op.forward(torch.from_numpy(img).to(dev).contiguous())
    torch.cuda.synchronize()
    out = out.double().cpu().numpy()
    rows = list(range(0, 16)) + list(range(1016, 1032)) + list(range(2032, 2048))
    for b in range(B):
        ref = np.concatenate([oracle.conv2d_same(img[b, 0].astype(np.float64), g, rows=(r0, r0 + 16))
                              for r0 in (0, 1016, 2032)])
        l2, mx = oracle.err_l2max(out[b, 0, rows], ref)
        print(f"2048^2 image {b}: rel_l2={l2:.2e} rel_max={mx:.2e}")
        assert l2 <= 1e-4 and mx <= 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["same", "valid_corr"])
def test_conv_forward_two_separates_paired_images(dev, mode):
    """Two images with different filters in one complex transform (spectra separated by
    X(k) +- conj X(-k)); each result against the direct FP64 oracle, tolerance as above."""
    import torch

    from fft_demo_amd.conv import Conv2d

    h, w, k, F = 200, 150, 7, 3
    filt = oracle.uniform(F * k * k, 41, -1.0, 1.0).reshape(F, 1, k, k)
    imgs = [oracle.uniform(h * w, 42 + i, 0.0, 1.0).reshape(h, w) for i in range(2)]
    op = Conv2d(1, 1, F, h, w, k, k, mode=mode)
    op.set_filters(torch.from_numpy(filt).to(device=dev, dtype=torch.float32).contiguous())
    xs = [torch.from_numpy(im).to(device=dev, dtype=torch.float32).contiguous() for im in imgs]
    for f0, f1 in [(2, 0), (1, 1)]:
        o0, o1 = op.forward_two(xs[0], xs[1], f0, f1)
        torch.cuda.synchronize()
        for im, f, o in [(imgs[0], f0, o0), (imgs[1], f1, o1)]:
            ref = oracle.conv2d_same(im, filt[f, 0]) if mode == "same" else oracle.corr2d_valid(im, filt[f, 0])
            got = o.double().cpu().numpy()
            l2 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            mx = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"forward_two {mode} f={f} rel_l2={l2:.2e} max={mx:.2e}")
            assert l2 <= 1e-4 and mx <= 1e-3


def _backward_case(dev, B, C, F, h, w, fh, fw, dtype, seed):
    import torch

    from fft_demo_amd.conv import Conv2d

    oh, ow = h - fh + 1, w - fw + 1
    x = oracle.uniform(B * C * h * w, seed, 0.0, 1.0).reshape(B, C, h, w)
    filt = oracle.uniform(F * C * fh * fw, seed + 1, -1.0, 1.0).reshape(F, C, fh, fw)
    delta = oracle.uniform(B * F * oh * ow, seed + 2, -1.0, 1.0).reshape(B, F, oh, ow)
    td = getattr(torch, dtype)
    op = Conv2d(B, C, F, h, w, fh, fw, mode="valid_corr")
    op.set_filters(torch.from_numpy(filt).to(device=dev, dtype=td).contiguous())
    xd = torch.from_numpy(x).to(device=dev, dtype=td).contiguous()
    dd = torch.from_numpy(delta).to(device=dev, dtype=td).contiguous()
    y = op.forward(xd)
    gi, gw = op.backward(dd, xd)
    torch.cuda.synchronize()
    return x, filt, delta, y.double().cpu().numpy(), gi.double().cpu().numpy(), gw.double().cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_conv_backward_matches_direct_gradients(dev, dtype):
    """Gradients of the 'valid' correlation layer (the reference's conv2d_backward,
    tensorcore_fft_conv2d.cpp:364-442) against the direct FP64 definitions: grad_w =
    sum_b corr_valid(in[b][c], delta[b][f]); grad_in = sum_f conv_full(delta[b][f], w[f][c]).
    Odd batch (the last image pair has an empty half). Tolerance rel-L2 <= 1e-4 per map."""
    B, C, F, h, w, fh, fw = 3, 2, 2, 40, 52, 5, 7
    x, filt, delta, _, gi, gw = _backward_case(dev, B, C, F, h, w, fh, fw, dtype, 61)
    for f in range(F):
        for c in range(C):
            ref = sum(oracle.corr2d_valid(x[b, c], delta[b, f]) for b in range(B))
            l2 = np.linalg.norm(gw[f, c] - ref) / np.linalg.norm(ref)
            print(f"grad_w f={f} c={c} rel_l2={l2:.2e}")
            assert l2 <= 1e-4
    for b in range(B):
        for c in range(C):
            ref = sum(oracle.conv2d_full(delta[b, f], filt[f, c]) for f in range(F))
            l2 = np.linalg.norm(gi[b, c] - ref) / np.linalg.norm(ref)
            print(f"grad_in b={b} c={c} rel_l2={l2:.2e}")
            assert l2 <= 1e-4


@pytest.mark.gpu
def test_conv_backward_adjoint_identities_large(dev):
    """Size-independent check at 1024 x 1024 with a 31 x 31 filter (pad 1024 = the image):
    <forward(x), delta> = <x, grad_in(delta)> = <w, grad_w(x, delta)>. The inner products
    cancel heavily, so the bound is 1e-4 of the Cauchy-Schwarz scale ||a|| ||b|| of each
    product (what a relative L2 error of 1e-4 in any one operand can move it by)."""
    B, C, F, h, w, k = 2, 1, 1, 1024, 1024, 31
    x, filt, delta, y, gi, gw = _backward_case(dev, B, C, F, h, w, k, k, "float32", 71)
    lhs = float(np.sum(y * delta))
    r_in = float(np.sum(x * gi))
    r_w = float(np.sum(filt * gw))
    scale = max(np.linalg.norm(y) * np.linalg.norm(delta), np.linalg.norm(x) * np.linalg.norm(gi),
                np.linalg.norm(filt) * np.linalg.norm(gw))
    print(f"<y,d>={lhs:.6e} <x,gi>={r_in:.6e} <w,gw>={r_w:.6e} scale={scale:.3e}")
    assert abs(r_in - lhs) <= 1e-4 * scale
    assert abs(r_w - lhs) <= 1e-4 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("B,C,F,h,w,fh,fw", [(3, 1, 1, 120, 136, 9, 7), (2, 2, 3, 100, 90, 5, 5)])
def test_conv_same_fp16_spectra(dev, B, C, F, h, w, fh, fw):
    """fp16 image spectra and intermediates (spec_fp16): one filter and channel (the image
    load and the filter product fused into the transforms' first passes) and several
    (separate channel-sum pass). Each pass boundary rounds to fp16, so the tolerance is
    rel-L2 <= 1e-3 and rel-max <= 1e-2 of each output map against the FP64 oracle."""
    import torch

    from fft_demo_amd.conv import Conv2d

    img = oracle.uniform(B * C * h * w, 71, 0.0, 1.0).reshape(B, C, h, w)
    filt = oracle.uniform(F * C * fh * fw, 72, -1.0, 1.0).reshape(F, C, fh, fw)
    op = Conv2d(B, C, F, h, w, fh, fw, mode="same", spec_fp16=True)
    op.set_filters(torch.from_numpy(filt).to(device=dev, dtype=torch.float32).contiguous())
    out = op.forward(torch.from_numpy(img).to(device=dev, dtype=torch.float32).contiguous())
    torch.cuda.synchronize()
    got = out.double().cpu().numpy()
    for b in range(B):
        for f in range(F):
            ref = sum(oracle.conv2d_same(img[b, c], filt[f, c]) for c in range(C))
            l2, mx = oracle.err_l2max(got[b, f], ref)
            print(f"spec_fp16 pad={op.pad} b={b} f={f} rel_l2={l2:.2e} rel_max={mx:.2e}")
            assert l2 <= 1e-3 and mx <= 1e-2, (b, f, l2, mx)


@pytest.mark.gpu
@pytest.mark.parametrize("B,h,w,fh,fw,pad", [(3, 120, 136, 9, 7, None), (2, 100, 90, 5, 5, (128, 128)),
                                              (5, 80, 130, 7, 7, (90, 144))])
def test_conv_same_fp32_spectra_fp16_boundaries(dev, B, h, w, fh, fw, pad):
    """fp32 image spectra with fp16 pass boundaries inside each transform (mid_fp16), three
    filters, each selected per call; padded 128 x 144 (odd part 9: image load and filter
    product fused into the first passes), 128 x 128 (no fused pass: the packed images and
    the pointwise product go through the transforms out of place) and 90 x 144 (the fused
    filter product's radix-9 pass has 16 x 90 = 1440 lines per transform, not a whole number
    of the 64-line groups its waves take, and three pairs per launch). Tolerance as for
    spec_fp16: rel-L2 <= 1e-3 and rel-max <= 1e-2 of each output map. A call for all filters
    at once is refused (one filter per call)."""
    import torch

    from fft_demo_amd.conv import Conv2d

    F = 3
    img = oracle.uniform(B * h * w, 91, 0.0, 1.0).reshape(B, 1, h, w)
    filt = oracle.uniform(F * fh * fw, 92, -1.0, 1.0).reshape(F, 1, fh, fw)
    op = Conv2d(B, 1, F, h, w, fh, fw, mode="same", pad=pad, mid_fp16=True)
    op.set_filters(torch.from_numpy(filt).to(device=dev, dtype=torch.float32).contiguous())
    x = torch.from_numpy(img).to(device=dev, dtype=torch.float32).contiguous()
    for f in (2, 0):
        got = op.forward(x, filter=f).double().cpu().numpy()
        for b in range(B):
            l2, mx = oracle.err_l2max(got[b, 0], oracle.conv2d_same(img[b, 0], filt[f, 0]))
            print(f"mid_fp16 pad={op.pad} b={b} f={f} rel_l2={l2:.2e} rel_max={mx:.2e}")
            assert l2 <= 1e-3 and mx <= 1e-2, (b, f, l2, mx)
    with pytest.raises(RuntimeError, match="one filter per call"):
        op.forward(x)
    op.close()


@pytest.mark.gpu
@pytest.mark.parametrize("pad", [None, (256, 256)])
def test_conv_forward_two_fp32_spectra_fp16_boundaries(dev, pad):
    """forward_two on a mid_fp16 operator (fused and unfused first passes) against the
    oracle, tolerance as above; its backward pass is refused."""
    import torch

    from fft_demo_amd.conv import Conv2d

    h, w, k, F = 200, 150, 7, 3
    filt = oracle.uniform(F * k * k, 93, -1.0, 1.0).reshape(F, 1, k, k)
    imgs = [oracle.uniform(h * w, 94 + i, 0.0, 1.0).reshape(h, w) for i in range(2)]
    op = Conv2d(1, 1, F, h, w, k, k, mode="valid_corr", pad=pad, mid_fp16=True)
    op.set_filters(torch.from_numpy(filt).to(device=dev, dtype=torch.float32).contiguous())
    xs = [torch.from_numpy(im).to(device=dev, dtype=torch.float32).contiguous() for im in imgs]
    o0, o1 = op.forward_two(xs[0], xs[1], 1, 2)
    torch.cuda.synchronize()
    for im, f, o in [(imgs[0], 1, o0), (imgs[1], 2, o1)]:
        l2, mx = oracle.err_l2max(o.double().cpu().numpy(), oracle.corr2d_valid(im, filt[f, 0]))
        print(f"forward_two mid_fp16 pad={op.pad} f={f} rel_l2={l2:.2e} rel_max={mx:.2e}")
        assert l2 <= 1e-3 and mx <= 1e-2
    with pytest.raises(RuntimeError, match="fp32 spectra"):
        op.backward(torch.zeros(1, F, h - k + 1, w - k + 1, device=dev), xs[0].reshape(1, 1, h, w))
    op.close()


@pytest.mark.gpu
@pytest.mark.parametrize("B,C,F,h,w,fh,fw,mode", [(1, 1, 16, 256, 256, 13, 13, "valid_corr"),
                                                  (3, 2, 2, 64, 80, 7, 5, "same")])
def test_conv_reference_method(dev, B, C, F, h, w, fh, fw, mode):
    """The reference's method (reference_method=True; tensorcore_fft_conv2d_kernel.cu:1033-1150):
    unscaled fp16 transforms, fp16 image spectra, channel sum in fp32, each output map
    quantised to fp16 with its power-of-ten scale, dequantised at the readout. First the
    reference's accuracy_conv shape (main.cpp:320-352: 256^2, 16 filters of 13^2, 'valid'
    correlation), then several channels and an odd batch in 'same' mode, against the FP64
    oracle; the fp16 spectra bound the error (rel-L2 <= 5e-3, rel-max <= 5e-2 of each map).
    forward_two and the backward pass are refused in this mode."""
    import torch

    from fft_demo_amd.conv import Conv2d

    img = oracle.uniform(B * C * h * w, 101, 0.0, 1.0).reshape(B, C, h, w)
    filt = oracle.uniform(F * C * fh * fw, 102, -1.0, 1.0).reshape(F, C, fh, fw)
    op = Conv2d(B, C, F, h, w, fh, fw, mode=mode, reference_method=True)
    op.set_filters(torch.from_numpy(filt).to(device=dev, dtype=torch.float32).contiguous())
    x = torch.from_numpy(img).to(device=dev, dtype=torch.float32).contiguous()
    got = op.forward(x).double().cpu().numpy()
    worst = (0.0, 0.0)
    for b in range(B):
        for f in range(F):
            if mode == "same":
                ref = sum(oracle.conv2d_same(img[b, c], filt[f, c]) for c in range(C))
            else:
                ref = sum(oracle.corr2d_valid(img[b, c], filt[f, c]) for c in range(C))
            l2, mx = oracle.err_l2max(got[b, f], ref)
            worst = (max(worst[0], l2), max(worst[1], mx))
            assert l2 <= 5e-3 and mx <= 5e-2, (b, f, l2, mx)
    print(f"reference method {mode} {h}x{w} pad={op.pad} C={C} F={F}: worst rel_l2={worst[0]:.2e} "
          f"rel_max={worst[1]:.2e}")
    if C == 1:
        with pytest.raises(RuntimeError, match="reference-method"):
            op.forward_two(x[0, 0], x[0, 0], 0, 0)
    if mode == "valid_corr":
        with pytest.raises(RuntimeError, match="fp32 spectra"):
            op.backward(torch.zeros(B, F, h - fh + 1, w - fw + 1, device=dev), x)
    op.close()


@pytest.mark.gpu
def test_conv_forward_two_fp16_spectra_and_backward_refused(dev):
    """forward_two with fp16 spectra against the oracle (tolerance as for spec_fp16 above);
    the backward pass of such an operator is refused (it needs fp32 spectra)."""
    import torch

    from fft_demo_amd.conv import Conv2d

    h, w, k, F = 200, 150, 7, 3
    filt = oracle.uniform(F * k * k, 81, -1.0, 1.0).reshape(F, 1, k, k)
    imgs = [oracle.uniform(h * w, 82 + i, 0.0, 1.0).reshape(h, w) for i in range(2)]
    op = Conv2d(1, 1, F, h, w, k, k, mode="valid_corr", spec_fp16=True)
    op.set_filters(torch.from_numpy(filt).to(device=dev, dtype=torch.float32).contiguous())
    xs = [torch.from_numpy(im).to(device=dev, dtype=torch.float32).contiguous() for im in imgs]
    o0, o1 = op.forward_two(xs[0], xs[1], 2, 0)
    torch.cuda.synchronize()
    for im, f, o in [(imgs[0], 2, o0), (imgs[1], 0, o1)]:
        l2, mx = oracle.err_l2max(o.double().cpu().numpy(), oracle.corr2d_valid(im, filt[f, 0]))
        print(f"forward_two spec_fp16 f={f} rel_l2={l2:.2e} rel_max={mx:.2e}")
        assert l2 <= 1e-3 and mx <= 1e-2
    with pytest.raises(RuntimeError, match="fp32 spectra"):
        op.backward(torch.zeros(1, F, h - k + 1, w - k + 1, device=dev), xs[0].reshape(1, 1, h, w))
