"""Host-side checks of the conv operator's argument validation (no GPU): the error codes
and the message every C entry point leaves in mfft_last_error()."""
import ctypes

ERR_INVALID = -1  # MFFT_ERR_INVALID in include/mfft.h
ERR_UNSUPPORTED = -2  # MFFT_ERR_UNSUPPORTED


def test_conv_create_errors_reach_last_error():
    from fft_demo_amd import _lib

    lib = _lib.load()
    d = _lib.ConvDesc()
    assert lib.mfft_conv2d_desc_default(ctypes.byref(d)) == 0
    assert d.spec_fp16 == 0 and d.precision == _lib.PRECISIONS["split_fp16"]
    h = ctypes.c_void_p()
    d.h = 0
    assert lib.mfft_conv2d_create(ctypes.byref(h), ctypes.byref(d)) == ERR_INVALID
    assert b"bad sizes" in lib.mfft_last_error()
    d.h, d.w, d.fh, d.fw = 32, 32, 3, 3
    d.spec_fp16 = 4
    assert lib.mfft_conv2d_create(ctypes.byref(h), ctypes.byref(d)) == ERR_INVALID
    assert b"spec_fp16" in lib.mfft_last_error()
    # fp32 spectra with fp16 pass boundaries: single-channel operators only
    d.spec_fp16 = _lib.CONV_SPEC_FP32_MID_FP16
    d.channels = 2
    assert lib.mfft_conv2d_create(ctypes.byref(h), ctypes.byref(d)) == ERR_UNSUPPORTED
    assert b"single-channel" in lib.mfft_last_error()
    d.channels = 1
    d.spec_fp16 = 0
    d.mode = 7
    assert lib.mfft_conv2d_create(ctypes.byref(h), ctypes.byref(d)) == ERR_INVALID
    assert b"mode" in lib.mfft_last_error()
    assert not h.value
