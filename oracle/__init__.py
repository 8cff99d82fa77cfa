"""FP64 CPU oracle (ctypes binding of oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY. Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker. The product package
(fft_demo_amd) never imports it.

Each function follows the reference (yuelingfan/fft_demo):
  gen_rev            tcfft_half.cu:1351-1374
  fft / fft2d        main.cpp:68-94 (FFTW FORWARD convention, unnormalised)
  corr2d_valid       main.cpp:97-116
  conv2d_full        tensorcore_fft_conv2d.cpp:400-442 (input gradient of the 'valid' layer)
  err_capped         main.cpp:118-134
  fft_batch_loop     main.cpp:68-80 (the reference's CPU harness shape)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.c_int64


def build(force: bool = False) -> None:
    """Compile liboracle.so with gcc (the oracle is plain C)."""
    args = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.oracle_gen_rev.argtypes = [ctypes.c_int, _ip, ctypes.c_int, _ip]
    lib.oracle_gen_rev.restype = ctypes.c_int
    lib.oracle_fill_uniform.argtypes = [_dp, _i64, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
    lib.oracle_splitmix64.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64), _i64]
    lib.oracle_round_fp16.argtypes = [_dp, _i64]
    lib.oracle_round_bf16.argtypes = [_dp, _i64]
    lib.oracle_dft.argtypes = [_dp, _dp, _i64, _i64, ctypes.c_int]
    lib.oracle_fft.argtypes = [_dp, _dp, _i64, _i64, ctypes.c_int, ctypes.c_int]
    lib.oracle_fft.restype = ctypes.c_int
    lib.oracle_fft2d.argtypes = [_dp, _dp, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int]
    lib.oracle_fft2d.restype = ctypes.c_int
    lib.oracle_fft_batch_loop.argtypes = [_dp, _dp, _i64, _i64]
    lib.oracle_fft_batch_loop.restype = ctypes.c_int
    lib.oracle_corr2d_valid.argtypes = [_dp, _dp, _dp] + [ctypes.c_int] * 4
    lib.oracle_conv2d_same.argtypes = [_dp, _dp, _dp] + [ctypes.c_int] * 7
    lib.oracle_conv2d_full.argtypes = [_dp, _dp, _dp] + [ctypes.c_int] * 4
    lib.oracle_err_capped.argtypes = [_dp, _dp, _i64, _i64]
    lib.oracle_err_capped.restype = ctypes.c_double
    lib.oracle_err_l2max.argtypes = [_dp, _dp, _i64, _dp, _dp]
    _lib = lib
    return lib


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _as_interleaved(x) -> np.ndarray:
    x = np.asarray(x)
    if np.iscomplexobj(x):
        return _f64(np.stack([x.real, x.imag], axis=-1))
    return _f64(x)


def _to_complex(a: np.ndarray) -> np.ndarray:
    return a[..., 0] + 1j * a[..., 1]


def gen_rev(n: int, radices) -> np.ndarray:
    """Digit-reversal table (tcfft_half.cu:1351-1374): position i holds element rev[i]."""
    rad = np.ascontiguousarray(radices, dtype=np.int32)
    out = np.empty(n, dtype=np.int32)
    rc = _load().oracle_gen_rev(int(n), rad.ctypes.data_as(_ip), len(rad), out.ctypes.data_as(_ip))
    if rc != 0:
        raise ValueError(f"radices {list(radices)} do not multiply to {n}")
    return out


def splitmix64(seed: int, count: int) -> np.ndarray:
    """The raw SplitMix64 stream (64-bit outputs) the uniform generator draws from."""
    out = np.empty(int(count), dtype=np.uint64)
    _load().oracle_splitmix64(ctypes.c_uint64(seed & (2**64 - 1)), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                              int(count))
    return out


def uniform(count: int, seed: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """SplitMix64 uniform stream: value = lo + (hi - lo) * ((x >> 11) * 2^-53)."""
    out = np.empty(int(count), dtype=np.float64)
    _load().oracle_fill_uniform(_ptr(out), int(count), ctypes.c_uint64(seed & (2**64 - 1)), lo, hi)
    return out


def round_fp16(a) -> np.ndarray:
    out = _f64(a).copy()
    _load().oracle_round_fp16(_ptr(out), out.size)
    return out


def round_bf16(a) -> np.ndarray:
    out = _f64(a).copy()
    _load().oracle_round_bf16(_ptr(out), out.size)
    return out


def dft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """Brute-force FP64 DFT along the last axis of a complex array."""
    x = np.asarray(x, dtype=np.complex128)
    n = x.shape[-1]
    batch = x.size // n if n else 0
    a = _as_interleaved(x.reshape(batch, n))
    out = np.empty_like(a)
    _load().oracle_dft(_ptr(a), _ptr(out), n, batch, int(sign))
    return _to_complex(out).reshape(x.shape)


def fft(x: np.ndarray, sign: int = -1, threads: int = 0) -> np.ndarray:
    """FP64 mixed-radix FFT along the last axis (unnormalised, sign -1 forward)."""
    x = np.asarray(x, dtype=np.complex128)
    n = x.shape[-1]
    batch = x.size // n
    a = _as_interleaved(x.reshape(batch, n))
    out = np.empty_like(a)
    if _load().oracle_fft(_ptr(a), _ptr(out), n, batch, int(sign), int(threads)):
        raise RuntimeError("oracle_fft failed")
    return _to_complex(out).reshape(x.shape)


def fft2d(x: np.ndarray, sign: int = -1, threads: int = 0) -> np.ndarray:
    """FP64 2D FFT over the last two axes [..., nx, ny] (y contiguous)."""
    x = np.asarray(x, dtype=np.complex128)
    nx, ny = x.shape[-2], x.shape[-1]
    batch = x.size // (nx * ny)
    a = _as_interleaved(x.reshape(batch, nx, ny))
    out = np.empty_like(a)
    if _load().oracle_fft2d(_ptr(a), _ptr(out), nx, ny, batch, int(sign), int(threads)):
        raise RuntimeError("oracle_fft2d failed")
    return _to_complex(out).reshape(x.shape)


def fft_batch_loop(data_interleaved: np.ndarray, n: int, batch: int) -> np.ndarray:
    """The reference's CPU harness shape (main.cpp:68-80), single thread."""
    a = _f64(data_interleaved)
    out = np.empty_like(a)
    if _load().oracle_fft_batch_loop(_ptr(a), _ptr(out), n, batch):
        raise RuntimeError("oracle_fft_batch_loop failed")
    return out


def corr2d_valid(signal: np.ndarray, filt: np.ndarray) -> np.ndarray:
    """'valid' cross-correlation (main.cpp:97-116)."""
    s = _f64(signal)
    f = _f64(filt)
    sy, sx = s.shape
    fy, fx = f.shape
    out = np.empty((sy - fy + 1, sx - fx + 1), dtype=np.float64)
    _load().oracle_corr2d_valid(_ptr(s), _ptr(f), _ptr(out), sy, sx, fy, fx)
    return out


def conv2d_full(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Direct FP64 'full' convolution (the input gradient of the 'valid' correlation layer,
    tensorcore_fft_conv2d.cpp:400-442)."""
    x = _f64(a)
    y = _f64(b)
    ah, aw = x.shape
    bh, bw = y.shape
    out = np.empty((ah + bh - 1, aw + bw - 1), dtype=np.float64)
    _load().oracle_conv2d_full(_ptr(x), _ptr(y), _ptr(out), ah, aw, bh, bw)
    return out


def conv2d_same(image: np.ndarray, filt: np.ndarray, rows=None, threads: int = 0) -> np.ndarray:
    """Direct FP64 'same' convolution (filter flipped, centred, zero boundary)."""
    im = _f64(image)
    f = _f64(filt)
    h, w = im.shape
    fh, fw = f.shape
    r0, r1 = (0, h) if rows is None else rows
    out = np.empty((r1 - r0, w), dtype=np.float64)
    _load().oracle_conv2d_same(_ptr(im), _ptr(f), _ptr(out), h, w, fh, fw, r0, r1, int(threads))
    return out


def err_capped(tested, standard) -> float:
    """Reference metric (main.cpp:118-134) on complex arrays of shape [batch, n]."""
    t = _as_interleaved(tested)
    s = _as_interleaved(standard)
    n = np.asarray(tested).shape[-1]
    batch = np.asarray(tested).size // n
    return float(_load().oracle_err_capped(_ptr(t), _ptr(s), n, batch))


def err_l2max(tested, standard):
    """(rel_l2, rel_max) over all real components."""
    t = _as_interleaved(tested).ravel()
    s = _as_interleaved(standard).ravel()
    a = ctypes.c_double()
    b = ctypes.c_double()
    _load().oracle_err_l2max(_ptr(t), _ptr(s), t.size, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value
