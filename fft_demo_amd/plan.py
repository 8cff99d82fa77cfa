"""Python handle over the native plan/exec C ABI (include/mfft.h).

Device buffers are torch tensors (torch owns HIP memory and streams here); every
transform runs in the HIP kernels of libmfft.so. Complex data is interleaved: an fp16
complex buffer is a torch.float16 tensor whose last dimension is 2 (re, im).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

_TORCH_DTYPE = {}


def _torch():
    import torch

    if not _TORCH_DTYPE:
        _TORCH_DTYPE.update({L.FMT_C16: torch.float16, L.FMT_CBF16: torch.bfloat16, L.FMT_C32: torch.float32})
    return torch


def fmt_dtype(fmt: int):
    _torch()
    return _TORCH_DTYPE[fmt]


def _enum(table: dict, v):
    return table[v] if isinstance(v, str) else int(v)


def make_desc(precision="split_fp16", direction=L.FORWARD, scale=None, in_fmt=None, out_fmt=None,
              in_order=L.ORDER_NATURAL, workspace_limit=0, split_hint=0, mid_fmt=None) -> L.Desc:
    lib = L.load()
    d = L.Desc()
    L.check(lib.mfft_desc_default(ctypes.byref(d)), "mfft_desc_default")
    d.precision = _enum(L.PRECISIONS, precision)
    d.direction = int(direction)
    if scale is None:
        scale = L.SCALE_STAGE if d.precision == L.PREC_SPLIT_FP16 else L.SCALE_NONE
    d.scale = int(scale)
    default_fmt = {L.PREC_FP16: L.FMT_C16, L.PREC_BF16: L.FMT_CBF16, L.PREC_SPLIT_FP16: L.FMT_C16,
                   L.PREC_FP32: L.FMT_C32}[d.precision]
    d.in_fmt = default_fmt if in_fmt is None else _enum(L.FORMATS, in_fmt)
    d.out_fmt = default_fmt if out_fmt is None else _enum(L.FORMATS, out_fmt)
    d.in_order = int(in_order)
    d.workspace_limit = int(workspace_limit)
    d.split_hint = int(split_hint)
    d.mid_fmt = -1 if mid_fmt is None else _enum(L.FORMATS, mid_fmt)
    return d


class Plan:
    """A 1D (shape=(n,)) or 2D (shape=(nx, ny)) batched complex FFT plan."""

    def __init__(self, shape, batch: int, **desc):
        lib = L.load()
        self.shape = tuple(int(s) for s in shape)
        self.batch = int(batch)
        self.desc = make_desc(**desc)
        h = ctypes.c_void_p()
        if len(self.shape) == 1:
            rc = lib.mfft_plan_create_1d(ctypes.byref(h), self.shape[0], self.batch, ctypes.byref(self.desc))
        elif len(self.shape) == 2:
            rc = lib.mfft_plan_create_2d(ctypes.byref(h), self.shape[0], self.shape[1], self.batch,
                                         ctypes.byref(self.desc))
        else:
            raise ValueError("shape must have 1 or 2 dimensions")
        L.check(rc, "mfft_plan_create")
        self._h = h
        inf = L.Info()
        L.check(lib.mfft_plan_info(self._h, ctypes.byref(inf)), "mfft_plan_info")
        self.info = inf.as_dict()
        self.out_scale = inf.out_scale

    @property
    def n(self) -> int:
        return int(np.prod(self.shape))

    def in_dtype(self):
        return fmt_dtype(self.desc.in_fmt)

    def out_dtype(self):
        return fmt_dtype(self.desc.out_fmt)

    def empty_in(self, device="cuda"):
        torch = _torch()
        return torch.empty((self.batch, *self.shape, 2), dtype=self.in_dtype(), device=device)

    def empty_out(self, device="cuda"):
        torch = _torch()
        return torch.empty((self.batch, *self.shape, 2), dtype=self.out_dtype(), device=device)

    def _check_buf(self, t, dtype, what):
        if t.dtype != dtype:
            raise TypeError(f"{what} dtype {t.dtype} != {dtype}")
        if not t.is_cuda:
            raise ValueError(f"{what} must be a device tensor")
        if not t.is_contiguous():
            raise ValueError(f"{what} must be contiguous")
        if t.numel() != self.batch * self.n * 2:
            raise ValueError(f"{what} has {t.numel()} values, expected {self.batch * self.n * 2}")

    def exec(self, x, out=None, stream=None):
        """Run the transform on device tensors; returns `out`."""
        torch = _torch()
        self._check_buf(x, self.in_dtype(), "input")
        if out is None:
            out = self.empty_out(device=x.device)
        self._check_buf(out, self.out_dtype(), "output")
        s = stream if stream is not None else torch.cuda.current_stream(x.device)
        rc = L.load().mfft_exec(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                ctypes.c_void_p(s.cuda_stream))
        L.check(rc, "mfft_exec")
        return out

    def exec_ptr(self, in_ptr: int, out_ptr: int, stream_handle: int = 0) -> None:
        rc = L.load().mfft_exec(self._h, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr),
                                ctypes.c_void_p(stream_handle))
        L.check(rc, "mfft_exec")

    def input_permutation(self, axis: int = 0) -> np.ndarray:
        n = self.shape[0] if len(self.shape) == 1 else self.shape[axis]
        out = np.empty(n, dtype=np.int32)
        L.check(L.load().mfft_input_permutation(self._h, axis, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n),
                "mfft_input_permutation")
        return out

    def close(self):
        if getattr(self, "_h", None):
            L.load().mfft_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
