"""fft_demo_amd -- MI355X (gfx950) matrix-core FFT and FFT-convolution.

The compute path is the native library lib/libmfft.so (hand-written HIP kernels and a
C++ planner behind the C ABI in include/mfft.h). Python modules here are thin handles:
  plan     -- Plan (1D/2D batched FFT) over mfft_plan_create_* / mfft_exec
  conv     -- Conv2d (FFT convolution operator, forward and backward) over mfft_conv2d_*
  serve    -- Server (continuous-batching request loop) over mfft_server_*
  tune     -- autotuner and its cache over mfft_autotune / mfft_tune_*
  hiptime  -- HIP events on the exec stream (timing)
The reference's own entry points (tcfftCreate/tcfftExec/setup/doit/finalize, the 2D ones and
fft_conv2d_half) are C/C++ headers in include/, for C++ callers.
"""
from ._lib import (FORWARD, INVERSE, SCALE_NONE, SCALE_STAGE, ORDER_NATURAL, ORDER_DIGIT_REVERSED,
                   PREC_FP16, PREC_BF16, PREC_SPLIT_FP16, PREC_FP32, FMT_C16, FMT_CBF16, FMT_C32,
                   MfftError, load)
from .plan import Plan, make_desc

__all__ = ["Plan", "make_desc", "load", "MfftError", "FORWARD", "INVERSE", "SCALE_NONE", "SCALE_STAGE",
           "ORDER_NATURAL", "ORDER_DIGIT_REVERSED", "PREC_FP16", "PREC_BF16", "PREC_SPLIT_FP16", "PREC_FP32",
           "FMT_C16", "FMT_CBF16", "FMT_C32"]
