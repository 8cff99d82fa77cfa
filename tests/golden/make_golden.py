"""Generates the committed golden fixtures under tests/golden/ (run in the build container).

Sources:
  rev_tables.json   the reference's own digit-reversal tables rev_256 / rev_512, read as
                    text (data) from tensorcore_fft_conv2d.h:72-120 of the reference.
  fft_vectors.npz   FP64 FFT vectors from numpy.fft (pocketfft, an independent library),
                    inputs from the SplitMix64 stream (oracle.uniform).
  conv_vectors.npz  FP64 'same' convolution and 'valid' correlation from scipy.signal.

Usage: python tests/golden/make_golden.py [/path/to/reference]
"""
from __future__ import annotations

import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def rev_tables(ref_dir: str) -> dict:
    text = open(os.path.join(ref_dir, "tensorcore_fft_conv2d.h")).read()
    out = {}
    for name in ("rev_256", "rev_512"):
        m = re.search(name + r"\[1024\]\s*=\s*\{([^}]*)\}", text)
        vals = [int(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
        out[name] = vals
    return out


def cplx(shape, seed):
    v = oracle.uniform(int(np.prod(shape)) * 2, seed).reshape(*shape, 2)
    return v[..., 0] + 1j * v[..., 1]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    if os.path.isdir(ref):
        with open(os.path.join(HERE, "rev_tables.json"), "w") as fh:
            json.dump(rev_tables(ref), fh)
    vec = {}
    for n, b, seed in [(16, 3, 1), (256, 2, 2), (512, 2, 3), (1024, 2, 4), (4096, 1, 5), (2176, 1, 6), (17, 2, 7)]:
        x = cplx((b, n), seed)
        vec[f"x_{n}"] = x
        vec[f"fwd_{n}"] = np.fft.fft(x, axis=-1)
        vec[f"inv_{n}"] = np.fft.ifft(x, axis=-1) * n
    x2 = cplx((2, 32, 64), 8)
    vec["x2d_32x64"] = x2
    vec["fwd2d_32x64"] = np.fft.fft2(x2)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **vec)

    from scipy import signal

    conv = {}
    img = oracle.uniform(64 * 64, 9, 0.0, 1.0).reshape(64, 64)
    for k in (3, 7, 31):
        f = oracle.uniform(k * k, 10 + k, 0.0, 1.0).reshape(k, k)
        conv[f"img"] = img
        conv[f"filt_{k}"] = f
        conv[f"same_{k}"] = signal.convolve2d(img, f, mode="same")
        conv[f"corr_valid_{k}"] = signal.correlate2d(img, f, mode="valid")
    np.savez_compressed(os.path.join(HERE, "conv_vectors.npz"), **conv)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
