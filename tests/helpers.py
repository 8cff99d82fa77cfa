"""Shared test helpers: seeded inputs, device upload/download and error metrics."""
from __future__ import annotations

import numpy as np

import oracle


def complex_input(shape, seed: int, fmt: str = "c16", lo=-1.0, hi=1.0) -> np.ndarray:
    """Interleaved re/im uniform[lo, hi] from the SplitMix64 stream, rounded to the
    boundary format so input quantisation is not counted as error."""
    count = int(np.prod(shape)) * 2
    v = oracle.uniform(count, seed, lo, hi)
    if fmt == "c16":
        v = oracle.round_fp16(v)
    elif fmt == "cbf16":
        v = oracle.round_bf16(v)
    v = v.reshape(*shape, 2)
    return v[..., 0] + 1j * v[..., 1]


def to_device(x: np.ndarray, dtype, device):
    import torch

    a = np.stack([x.real, x.imag], axis=-1)
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype).contiguous()


def from_device(t) -> np.ndarray:
    a = t.detach().float().cpu().numpy().astype(np.float64)
    return a[..., 0] + 1j * a[..., 1]


def rel_errors(y: np.ndarray, ref: np.ndarray):
    return oracle.err_l2max(y, ref)
