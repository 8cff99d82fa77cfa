"""HIP event timing through the HIP runtime's C API (libamdhip64), for timing every kernel
of an exec on the stream it runs on (mfft_exec_events)."""
from __future__ import annotations

import ctypes
import ctypes.util

_hip = None


def _load():
    global _hip
    if _hip is None:
        name = ctypes.util.find_library("amdhip64") or "libamdhip64.so"
        for cand in (name, "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64 not found")
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
    return _hip


class EventPool:
    """A pool of hipEvent_t handles (as a ctypes array usable by mfft_exec_events)."""

    def __init__(self, n: int):
        hip = _load()
        self.n = n
        self.arr = (ctypes.c_void_p * n)()
        for i in range(n):
            h = ctypes.c_void_p()
            rc = hip.hipEventCreate(ctypes.byref(h))
            if rc != 0:
                raise RuntimeError(f"hipEventCreate failed ({rc})")
            self.arr[i] = h

    def slice(self, start: int):
        return ctypes.cast(ctypes.byref(self.arr, start * ctypes.sizeof(ctypes.c_void_p)),
                           ctypes.POINTER(ctypes.c_void_p))

    def record(self, i: int, stream_handle: int) -> None:
        rc = _load().hipEventRecord(self.arr[i], ctypes.c_void_p(stream_handle))
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed ({rc})")

    def elapsed_ms(self, i: int, j: int) -> float:
        hip = _load()
        hip.hipEventSynchronize(self.arr[j])
        out = ctypes.c_float()
        rc = hip.hipEventElapsedTime(ctypes.byref(out), self.arr[i], self.arr[j])
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime failed ({rc})")
        return float(out.value)

    def close(self):
        hip = _load()
        for i in range(self.n):
            if self.arr[i]:
                hip.hipEventDestroy(self.arr[i])
                self.arr[i] = None
