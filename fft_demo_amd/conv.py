"""FFT-based 2D convolution operator: Python handle over mfft_conv2d_* (include/mfft.h).

Conv2d(batch, channels, filters, h, w, fh, fw, mode="same"|"valid_corr") computes
out[b][f] = sum_c in[b][c] (*) filt[f][c] in the HIP kernels of libmfft.so. Filter spectra
are computed once by set_filters() and stay resident on the device.
"""
from __future__ import annotations

import ctypes

from . import _lib as L

MODES = {"same": L.CONV_SAME, "valid_corr": L.CONV_VALID_CORR}


class Conv2d:
    def __init__(self, batch, channels, filters, h, w, fh, fw, mode="same", pad=None, precision="split_fp16",
                 workspace_limit=0, spec_fp16=False, mid_fp16=False, reference_method=False):
        """spec_fp16: fp16 image spectra and intermediates; mid_fp16: fp32 image spectra with
        fp16 pass boundaries inside each transform (one channel, one filter per forward call);
        reference_method: the reference's fp16 method with per-map power-of-ten quantisation
        of the product spectra. All three serve forward calls only."""
        lib = L.load()
        d = L.ConvDesc()
        lib.mfft_conv2d_desc_default(ctypes.byref(d))
        d.batch, d.channels, d.filters, d.h, d.w, d.fh, d.fw = batch, channels, filters, h, w, fh, fw
        d.mode = MODES[mode] if isinstance(mode, str) else int(mode)
        if pad is not None:
            d.pad_h, d.pad_w = (pad, pad) if isinstance(pad, int) else pad
        d.precision = L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        d.workspace_limit = workspace_limit
        if spec_fp16 + mid_fp16 + reference_method > 1:
            raise ValueError("spec_fp16, mid_fp16 and reference_method exclude each other")
        d.spec_fp16 = (L.CONV_SPEC_FP16 if spec_fp16 else L.CONV_SPEC_FP32_MID_FP16 if mid_fp16
                       else L.CONV_SPEC_REFERENCE if reference_method else L.CONV_SPEC_FP32)
        h_ = ctypes.c_void_p()
        L.check(lib.mfft_conv2d_create(ctypes.byref(h_), ctypes.byref(d)), "mfft_conv2d_create")
        self._h = h_
        self.desc = d
        inf = L.ConvInfo()
        L.check(lib.mfft_conv2d_info_get(self._h, ctypes.byref(inf)), "mfft_conv2d_info_get")
        self.pad = (inf.pad_h, inf.pad_w)
        self.out_shape = (batch, filters, inf.out_h, inf.out_w)
        self.device_bytes = inf.device_bytes
        self.spectrum_bytes = inf.spectrum_bytes

    @staticmethod
    def _fmt(t):
        import torch

        if t.dtype == torch.float32:
            return L.REAL_F32
        if t.dtype == torch.float64:
            return L.REAL_F64
        raise TypeError(f"unsupported dtype {t.dtype}")

    def set_filters(self, filt, stream=None):
        import torch

        d = self.desc
        assert filt.is_cuda and filt.is_contiguous() and filt.numel() == d.filters * d.channels * d.fh * d.fw
        s = stream if stream is not None else torch.cuda.current_stream(filt.device)
        L.check(L.load().mfft_conv2d_set_filters(self._h, ctypes.c_void_p(filt.data_ptr()), self._fmt(filt),
                                                  ctypes.c_void_p(s.cuda_stream)), "mfft_conv2d_set_filters")

    def forward(self, x, out=None, stream=None, filter=None):
        """All filters (out: [batch][filters][oh][ow]), or only filter index `filter`
        (mfft_conv2d_forward_select; out: [batch][1][oh][ow])."""
        import torch

        d = self.desc
        assert x.is_cuda and x.is_contiguous() and x.numel() == d.batch * d.channels * d.h * d.w
        if out is None:
            shape = self.out_shape if filter is None else (d.batch, 1) + self.out_shape[2:]
            out = torch.empty(shape, dtype=x.dtype, device=x.device)
        s = stream if stream is not None else torch.cuda.current_stream(x.device)
        lib = L.load()
        if filter is None:
            L.check(lib.mfft_conv2d_forward(self._h, ctypes.c_void_p(x.data_ptr()), self._fmt(x),
                                            ctypes.c_void_p(out.data_ptr()), self._fmt(out),
                                            ctypes.c_void_p(s.cuda_stream)), "mfft_conv2d_forward")
        else:
            L.check(lib.mfft_conv2d_forward_select(self._h, ctypes.c_void_p(x.data_ptr()), self._fmt(x),
                                                   ctypes.c_void_p(out.data_ptr()), self._fmt(out), int(filter),
                                                   ctypes.c_void_p(s.cuda_stream)), "mfft_conv2d_forward_select")
        return out

    def forward_two(self, x0, x1, f0, f1, out0=None, out1=None, stream=None):
        """Single-channel operators: x0 (*) filter f0 and x1 (*) filter f1 in one complex
        transform (x0, x1: [h][w] images of one dtype)."""
        import torch

        d = self.desc
        for x in (x0, x1):
            assert x.is_cuda and x.is_contiguous() and x.numel() == d.h * d.w and x.dtype == x0.dtype
        oh, ow = self.out_shape[2:]
        out0 = torch.empty((oh, ow), dtype=x0.dtype, device=x0.device) if out0 is None else out0
        out1 = torch.empty((oh, ow), dtype=x0.dtype, device=x0.device) if out1 is None else out1
        s = stream if stream is not None else torch.cuda.current_stream(x0.device)
        f = self._fmt(x0)
        L.check(L.load().mfft_conv2d_forward_two(self._h, ctypes.c_void_p(x0.data_ptr()), ctypes.c_void_p(x1.data_ptr()),
                                                  f, ctypes.c_void_p(out0.data_ptr()), ctypes.c_void_p(out1.data_ptr()),
                                                  f, int(f0), int(f1), ctypes.c_void_p(s.cuda_stream)),
                "mfft_conv2d_forward_two")
        return out0, out1

    def backward(self, delta, x=None, grad_in=True, grad_w=True, stream=None):
        """Gradients of the 'valid_corr' operator: delta = dL/dout [B][F][oh][ow], x = the
        forward input [B][C][h][w] (needed for grad_w). Returns (grad_in, grad_w), each
        None when not asked for."""
        import torch

        d = self.desc
        assert delta.is_cuda and delta.is_contiguous() and delta.numel() == d.batch * d.filters * self.out_shape[2] * \
            self.out_shape[3]
        gi = torch.empty((d.batch, d.channels, d.h, d.w), dtype=delta.dtype, device=delta.device) if grad_in else None
        gw = torch.empty((d.filters, d.channels, d.fh, d.fw), dtype=delta.dtype, device=delta.device) \
            if grad_w else None
        if grad_w:
            assert x is not None and x.is_cuda and x.is_contiguous() and x.dtype == delta.dtype
        s = stream if stream is not None else torch.cuda.current_stream(delta.device)
        ptr = (lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None)
        L.check(L.load().mfft_conv2d_backward(self._h, ptr(x), ptr(delta), self._fmt(delta), ptr(gi), ptr(gw),
                                               ctypes.c_void_p(s.cuda_stream)), "mfft_conv2d_backward")
        return gi, gw

    def close(self):
        if getattr(self, "_h", None):
            L.load().mfft_conv2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def good_size(n: int) -> int:
    return int(L.load().mfft_good_size(int(n)))
