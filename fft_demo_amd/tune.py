"""Schedule autotuner (mfft_autotune / mfft_tune_* in include/mfft.h).

autotune(shape, batch, **desc) times the planner's candidate schedules for one plan on the
current device -- the four-step split point and the batch chunk, never the precision --
and caches the fastest; Plan(shape, batch, split_hint=SPLIT_TUNED, ...) then uses it. The
cache is process-wide; save()/load() keep it in a text file across processes.
"""
from __future__ import annotations

import ctypes

from . import _lib as L
from .plan import make_desc

SPLIT_TUNED = L.SPLIT_TUNED


def _dims(shape):
    shape = tuple(int(s) for s in shape)
    if len(shape) == 1:
        return shape[0], 0
    if len(shape) == 2:
        return shape
    raise ValueError("shape must have 1 or 2 dimensions")


def _entry(e):
    return {"split_hint": e.split_hint, "workspace_limit": e.workspace_limit, "ms": e.ms,
            "default_ms": e.default_ms, "candidates": e.candidates}


def autotune(shape, batch, reps=5, **desc) -> dict:
    lib = L.load()
    nx, ny = _dims(shape)
    d = make_desc(**desc)
    e = L.TuneEntry()
    rc = lib.mfft_autotune(nx, ny, int(batch), ctypes.byref(d), int(reps), ctypes.byref(e))
    if rc != 0:
        raise RuntimeError(f"mfft_autotune: rc={rc}: {lib.mfft_tune_last_error().decode()}")
    return _entry(e)


def lookup(shape, batch, arch=None, **desc):
    """The cached schedule for this plan, or None."""
    lib = L.load()
    nx, ny = _dims(shape)
    d = make_desc(**desc)
    e = L.TuneEntry()
    rc = lib.mfft_tune_lookup(arch.encode() if arch else None, nx, ny, int(batch), ctypes.byref(d), ctypes.byref(e))
    return _entry(e) if rc == 0 else None


def save(path):
    lib = L.load()
    if lib.mfft_tune_cache_save(str(path).encode()) != 0:
        raise OSError(lib.mfft_tune_last_error().decode())


def load(path):
    lib = L.load()
    rc = lib.mfft_tune_cache_load(str(path).encode())
    if rc != 0:
        raise ValueError(lib.mfft_tune_last_error().decode())


def clear():
    L.load().mfft_tune_cache_clear()
