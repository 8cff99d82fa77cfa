"""ctypes binding of the native library fft_demo_amd/lib/libmfft.so (C ABI in include/mfft.h).

The native library is the product; this module only loads it and declares signatures.
It raises if the library is missing: there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmfft.so")

MFFT_OK = 0
ERRORS = {-1: "MFFT_ERR_INVALID", -2: "MFFT_ERR_UNSUPPORTED", -3: "MFFT_ERR_HIP", -4: "MFFT_ERR_NOMEM"}

PREC_FP16, PREC_BF16, PREC_SPLIT_FP16, PREC_FP32 = 0, 1, 2, 3
FMT_C16, FMT_CBF16, FMT_C32 = 0, 1, 2
FORWARD, INVERSE = -1, 1
SCALE_NONE, SCALE_STAGE = 0, 1
ORDER_NATURAL, ORDER_DIGIT_REVERSED = 0, 1

PRECISIONS = {"fp16": PREC_FP16, "bf16": PREC_BF16, "split_fp16": PREC_SPLIT_FP16, "fp32": PREC_FP32}
FORMATS = {"c16": FMT_C16, "cbf16": FMT_CBF16, "c32": FMT_C32}


class MfftError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed: {ERRORS.get(code, code)}: {msg}")
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int),
        ("direction", ctypes.c_int),
        ("scale", ctypes.c_int),
        ("in_fmt", ctypes.c_int),
        ("out_fmt", ctypes.c_int),
        ("in_order", ctypes.c_int),
        ("workspace_limit", ctypes.c_int64),
        ("split_hint", ctypes.c_int),
        ("mid_fmt", ctypes.c_int),
        ("external_workspace", ctypes.c_int),
    ]


KERNEL_NAMES = {0: "tile_pow2", 1: "tile_generic", 2: "fast_rows", 3: "fast_rows_tr", 4: "fast_cols", 6: "onchip", 7: "onchip_cols",
                5: "stream_small"}


class PassInfo(ctypes.Structure):
    _fields_ = [
        ("R", ctypes.c_int),
        ("lines_per_tile", ctypes.c_int),
        ("lane_lines", ctypes.c_int),
        ("nstages", ctypes.c_int),
        ("radix", ctypes.c_int * 8),
        ("tiles", ctypes.c_int64),
        ("generic", ctypes.c_int),
        ("in_elem_bytes", ctypes.c_int),
        ("out_elem_bytes", ctypes.c_int),
        ("points", ctypes.c_int64),
        ("kernel", ctypes.c_int),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("n", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("batch", ctypes.c_int64),
        ("npasses", ctypes.c_int),
        ("passes", PassInfo * 6),
        ("out_scale", ctypes.c_double),
        ("workspace_bytes", ctypes.c_int64),
        ("batch_chunk", ctypes.c_int64),
        ("launches_per_exec", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {
            "kind": self.kind,
            "n": self.n,
            "nx": self.nx,
            "ny": self.ny,
            "batch": self.batch,
            "out_scale": self.out_scale,
            "workspace_bytes": self.workspace_bytes,
            "batch_chunk": self.batch_chunk,
            "launches_per_exec": self.launches_per_exec,
            "passes": [
                {
                    "R": p.R,
                    "lines_per_tile": p.lines_per_tile,
                    "lane_lines": p.lane_lines,
                    "radix": list(p.radix[: p.nstages]),
                    "tiles": p.tiles,
                    "generic": bool(p.generic),
                    "in_elem_bytes": p.in_elem_bytes,
                    "out_elem_bytes": p.out_elem_bytes,
                    "kernel": KERNEL_NAMES.get(p.kernel, str(p.kernel)),
                    "points": p.points,
                }
                for p in self.passes[: self.npasses]
            ],
        }


_lib = None

# every symbol include/*.h declares, with its ctypes signature (restype, argtypes)
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.c_int64
SIGNATURES = {
    # include/mfft.h
    "mfft_desc_default": (ctypes.c_int, [ctypes.POINTER(Desc)]),
    "mfft_plan_create_1d": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _i64, ctypes.POINTER(Desc)]),
    "mfft_plan_create_2d": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _i64, ctypes.POINTER(Desc)]),
    "mfft_plan_destroy": (ctypes.c_int, [_vp]),
    "mfft_plan_info": (ctypes.c_int, [_vp, ctypes.POINTER(Info)]),
    "mfft_exec": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "mfft_exec_events": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(_vp), ctypes.c_int]),
    "mfft_exec_pass": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "mfft_input_permutation": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.c_int]),
    "mfft_graph_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, _vp, _vp, _vp]),
    "mfft_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "mfft_graph_destroy": (ctypes.c_int, [_vp]),
    "mfft_stage_convert": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, _i64, _i64, _vp, ctypes.c_int, _vp,
                                          ctypes.c_int, ctypes.c_double, _vp]),
    "mfft_last_error": (ctypes.c_char_p, []),
    "mfft_version": (ctypes.c_char_p, []),
}


class ConvDesc(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_int), ("channels", ctypes.c_int), ("filters", ctypes.c_int), ("h", ctypes.c_int),
                ("w", ctypes.c_int), ("fh", ctypes.c_int), ("fw", ctypes.c_int), ("mode", ctypes.c_int),
                ("pad_h", ctypes.c_int), ("pad_w", ctypes.c_int), ("precision", ctypes.c_int),
                ("spec_fp16", ctypes.c_int), ("workspace_limit", ctypes.c_int64)]


class ConvInfo(ctypes.Structure):
    _fields_ = [("pad_h", ctypes.c_int), ("pad_w", ctypes.c_int), ("out_h", ctypes.c_int), ("out_w", ctypes.c_int),
                ("device_bytes", ctypes.c_int64), ("spectrum_bytes", ctypes.c_int64)]


CONV_SAME, CONV_VALID_CORR = 0, 1
REAL_F32, REAL_F64 = 0, 1
CONV_SPEC_FP32, CONV_SPEC_FP16, CONV_SPEC_FP32_MID_FP16, CONV_SPEC_REFERENCE = 0, 1, 2, 3

SIGNATURES.update({
    "mfft_conv2d_desc_default": (ctypes.c_int, [ctypes.POINTER(ConvDesc)]),
    "mfft_conv2d_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(ConvDesc)]),
    "mfft_conv2d_set_filters": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    "mfft_conv2d_forward": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp]),
    "mfft_conv2d_forward_select": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "mfft_conv2d_forward_two": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, _vp]),
    "mfft_conv2d_backward": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "mfft_conv2d_info_get": (ctypes.c_int, [_vp, ctypes.POINTER(ConvInfo)]),
    "mfft_conv2d_destroy": (ctypes.c_int, [_vp]),
    "mfft_good_size": (ctypes.c_int, [ctypes.c_int]),
})


class BatchItem(ctypes.Structure):
    _fields_ = [("in_", ctypes.c_void_p), ("out", ctypes.c_void_p), ("n", ctypes.c_int), ("count", ctypes.c_int64)]


SIGNATURES.update({
    "mfft_batcher_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(Desc), ctypes.c_int]),
    "mfft_batcher_exec": (ctypes.c_int, [_vp, ctypes.POINTER(BatchItem), ctypes.c_int, _vp]),
    "mfft_batcher_destroy": (ctypes.c_int, [_vp]),
    "mfft_exec_ex": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _vp]),
    "mfft_exec_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp]),
})


class LaunchRecord(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_char * 64), ("bytes_read", ctypes.c_int64), ("bytes_written", ctypes.c_int64)]


SIGNATURES.update({
    "mfft_trace_begin": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(LaunchRecord), ctypes.c_int]),
    "mfft_trace_end": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
})


class TuneEntry(ctypes.Structure):
    _fields_ = [("split_hint", ctypes.c_int), ("workspace_limit", ctypes.c_int64), ("ms", ctypes.c_double),
                ("default_ms", ctypes.c_double), ("candidates", ctypes.c_int)]


SPLIT_TUNED = -1
SIGNATURES.update({
    "mfft_autotune": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(Desc), ctypes.c_int,
                                     ctypes.POINTER(TuneEntry)]),
    "mfft_tune_lookup": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                        ctypes.POINTER(Desc), ctypes.POINTER(TuneEntry)]),
    "mfft_tune_cache_save": (ctypes.c_int, [ctypes.c_char_p]),
    "mfft_tune_cache_load": (ctypes.c_int, [ctypes.c_char_p]),
    "mfft_tune_cache_clear": (ctypes.c_int, []),
    "mfft_tune_last_error": (ctypes.c_char_p, []),
})


class ServerDesc(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_int64), ("streams", ctypes.c_int), ("precision", ctypes.c_int),
                ("max_batch_1d", ctypes.c_int), ("max_batch_2d", ctypes.c_int), ("coalesce_max", ctypes.c_int),
                ("slab_bytes", ctypes.c_int64), ("plain", ctypes.c_int), ("mid_fmt", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("direction", ctypes.c_int), ("n", ctypes.c_int), ("nx", ctypes.c_int),
                ("ny", ctypes.c_int), ("batch", ctypes.c_int64), ("in_", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("filter_id", ctypes.c_int), ("tag", ctypes.c_int64)]


class Completion(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int64), ("status", ctypes.c_int), ("submit_us", ctypes.c_double),
                ("start_us", ctypes.c_double), ("end_us", ctypes.c_double), ("out_scale", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("coalesced", ctypes.c_int64), ("static_bytes", ctypes.c_int64), ("inflight_bytes", ctypes.c_int64),
                ("peak_bytes", ctypes.c_int64), ("paired", ctypes.c_int64), ("busy_us", ctypes.c_int64)]


REQ_FFT1D, REQ_FFT2D, REQ_CONV2D = 1, 2, 3
SIGNATURES.update({
    "mfft_server_desc_default": (ctypes.c_int, [ctypes.POINTER(ServerDesc)]),
    "mfft_server_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(ServerDesc)]),
    "mfft_server_add_filter": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                              ctypes.c_int]),
    "mfft_server_submit": (ctypes.c_int, [_vp, ctypes.POINTER(Request)]),
    "mfft_server_poll": (ctypes.c_int, [_vp, ctypes.POINTER(Completion), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "mfft_server_wait_idle": (ctypes.c_int, [_vp, ctypes.c_double]),
    "mfft_server_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ServerStats)]),
    "mfft_server_destroy": (ctypes.c_int, [_vp]),
    "mfft_server_last_error": (ctypes.c_char_p, []),
})


class TcfftHandle(ctypes.Structure):
    """struct tcfftHandle of include/tcfft_half.h (reference tcfft_half.h:7-18)."""
    _fields_ = [("N", ctypes.c_int), ("N_batch", ctypes.c_int), ("radices", ctypes.c_int * 9),
                ("n_radices", ctypes.c_int), ("mergings", ctypes.c_int * 3), ("n_mergings", ctypes.c_int),
                ("impl", ctypes.c_void_p)]


class TcfftHandle2d(ctypes.Structure):
    """struct tcfftHandle_2d of include/tcfft_half_2d.h (reference tcfft_half_2d.h:7-18)."""
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("N_batch", ctypes.c_int),
                ("radices_x", ctypes.c_int * 3), ("radices_y", ctypes.c_int * 3), ("n_radices_x", ctypes.c_int),
                ("n_radices_y", ctypes.c_int), ("mergings", ctypes.c_int * 2), ("impl", ctypes.c_void_p)]


_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES.update({
    # include/tcfft_half.h
    "tcfftCreate": (None, [ctypes.POINTER(TcfftHandle), ctypes.c_int, ctypes.c_int]),
    "tcfftExec": (None, [TcfftHandle, _vp]),
    "tcfftDestroy": (None, [ctypes.POINTER(TcfftHandle)]),
    "tcfftGetLastError": (ctypes.c_int, []),
    "setup": (None, [_dp, ctypes.c_int, ctypes.c_int]),
    "doit": (None, [ctypes.c_int]),
    "finalize": (None, [_dp, ctypes.c_int]),
    "setup_cu": (None, [_dp, ctypes.c_int, ctypes.c_int]),
    "doit_cu": (None, [ctypes.c_int]),
    "finalize_cu": (None, [_dp, ctypes.c_int, ctypes.c_int]),
    "setup_cu_2d": (None, [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "doit_cu_2d": (None, [ctypes.c_int]),
    "finalize_cu_2d": (None, [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    # include/tcfft_half_2d.h
    "tcfftCreate_2d": (None, [ctypes.POINTER(TcfftHandle2d), ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "tcfftExec_2d": (None, [TcfftHandle2d, _vp]),
    "tcfftDestroy_2d": (None, [ctypes.POINTER(TcfftHandle2d)]),
    "setup_2d": (None, [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "doit_2d": (None, [ctypes.c_int]),
    "finalize_2d": (None, [_dp, ctypes.c_int]),
})


def load(path: str | None = None):
    """Load libmfft.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"native library {p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int, where: str) -> None:
    if code != MFFT_OK:
        msg = load().mfft_last_error().decode(errors="replace")
        raise MfftError(code, where, msg)
