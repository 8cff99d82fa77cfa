"""GPU parity of the reference-named API (include/tcfft_half.h, tcfft_half_2d.h) and the
on-device staging kernel, called through the C ABI.

Reference contract (tcfft_half.cu:1177-1233, 1376-1413): tcfftExec transforms device
half2 data in place, input digit-reversed with gen_rev(plan.radices), output natural
order, unnormalised forward DFT. The precision here is fp16 matrix-core operands with
fp32 accumulation; tolerance rel-L2 <= 4e-3, rel-max <= 2e-2 vs the FP64 oracle on the
fp16-rounded input.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.helpers import complex_input, from_device, rel_errors, to_device

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    from fft_demo_amd import _lib as L

    return L, L.load()


# The reference's radix lists (tcfftCreate, tcfft_half.cu:1250-1330; radices default to 16,
# tcfft_half.h:10), which fix the digit-reversed input order its callers produce; sizes the
# reference has no case for take radix-16 digits then the tail (4096, 65536).
REF_RADICES = {
    256: [16, 16], 512: [16, 16, 2], 1024: [16, 16, 4], 4096: [16, 16, 16], 65536: [16, 16, 16, 16],
    131072: [16, 16, 2, 16, 16], 262144: [16, 16, 2, 16, 16, 2], 524288: [16, 16, 4, 16, 16, 2],
}


@pytest.mark.gpu
@pytest.mark.parametrize("n,batch", [(256, 4096), (512, 64), (1024, 32), (4096, 8), (65536, 4), (131072, 2),
                                     (262144, 2), (524288, 1)])
def test_tcfft_exec_digit_reversed_in_place(dev, n, batch):
    """Input permuted with gen_rev on the reference's own radix list for the size (not the
    handle's), as the reference's setup() does (tcfft_half.cu:1385-1391)."""
    import torch

    L, lib = _lib()
    h = L.TcfftHandle()
    lib.tcfftCreate(ctypes.byref(h), n, batch)
    assert lib.tcfftGetLastError() == 0, lib.mfft_last_error()
    radices = REF_RADICES[n]
    assert list(h.radices[: h.n_radices]) == radices
    assert list(h.radices[h.n_radices:]) == [16] * (9 - h.n_radices)
    x = complex_input((batch, n), 100 + n % 97)
    rev = oracle.gen_rev(n, radices)
    buf = to_device(x[:, rev], torch.float16, dev)  # setup()'s host permutation
    lib.tcfftExec(h, ctypes.c_void_p(buf.data_ptr()))
    assert lib.tcfftGetLastError() == 0
    torch.cuda.synchronize()
    y = from_device(buf)
    ref = oracle.fft(x)
    l2, mx = rel_errors(y, ref)
    print(f"tcfftExec n={n} b={batch} radices={radices} passes={h.n_mergings} rel_l2={l2:.2e} max={mx:.2e}")
    assert l2 <= 4e-3 and mx <= 2e-2
    lib.tcfftDestroy(ctypes.byref(h))


@pytest.mark.gpu
def test_harness_setup_doit_finalize(dev):
    """main.cpp:164-180 shape: natural-order FP64 host data in and out."""
    L, lib = _lib()
    n, b = 256, 8
    v = np.floor(oracle.uniform(2 * n * b, 42, 0, 10)) * 0.1  # the reference's 0.1*(rand()%10) values
    data = np.ascontiguousarray(v)
    res = np.zeros_like(data)
    dp = data.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    lib.setup(dp, n, b)
    lib.doit(1)
    lib.finalize(res.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), b)
    assert lib.tcfftGetLastError() == 0
    x = oracle.round_fp16(data).reshape(b, n, 2)
    ref = oracle.fft(x[..., 0] + 1j * x[..., 1])
    r = res.reshape(b, n, 2)
    l2, mx = rel_errors(r[..., 0] + 1j * r[..., 1], ref)
    capped = oracle.err_capped(r[..., 0] + 1j * r[..., 1], ref)
    print(f"harness rel_l2={l2:.2e} capped(main.cpp:118)={capped:.2e}")
    assert l2 <= 4e-3
    # comparator harness (plain fp32)
    lib.setup_cu(dp, n, b)
    lib.doit_cu(1)
    lib.finalize_cu(res.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, b)
    r = res.reshape(b, n, 2)
    ref32 = oracle.fft(data.reshape(b, n, 2)[..., 0] + 1j * data.reshape(b, n, 2)[..., 1])
    l2, _ = rel_errors(r[..., 0] + 1j * r[..., 1], ref32)
    assert l2 <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny", [(256, 256), (512, 256), (256, 1024), (1024, 1024)])
def test_tcfft_exec_2d_digit_reversed(dev, nx, ny):
    import torch

    L, lib = _lib()
    h = L.TcfftHandle2d()
    b = 2
    lib.tcfftCreate_2d(ctypes.byref(h), nx, ny, b)
    assert lib.tcfftGetLastError() == 0, lib.mfft_last_error()
    rx = oracle.gen_rev(nx, list(h.radices_x[: h.n_radices_x]))
    ry = oracle.gen_rev(ny, list(h.radices_y[: h.n_radices_y]))
    x = complex_input((b, nx, ny), 7)
    buf = to_device(x[:, rx][:, :, ry], torch.float16, dev)  # setup_2d (tcfft_half_2d.cu:1264-1271)
    lib.tcfftExec_2d(h, ctypes.c_void_p(buf.data_ptr()))
    torch.cuda.synchronize()
    assert lib.tcfftGetLastError() == 0
    y = from_device(buf)
    ref = oracle.fft2d(x)
    l2, mx = rel_errors(y, ref)
    print(f"tcfftExec_2d {nx}x{ny} rel_l2={l2:.2e} max={mx:.2e}")
    assert l2 <= 4e-3 and mx <= 2e-2
    lib.tcfftDestroy_2d(ctypes.byref(h))


@pytest.mark.gpu
def test_stage_convert_rounding_and_permutation(dev):
    import torch

    L, lib = _lib()
    n, b = 512, 3
    v = np.concatenate([oracle.uniform(2 * n * b - 6, 5, -70000, 70000), [65519.9, 65520.0, 1e-8, -3e-5, 2.0 ** -25, 0.0]])
    src = torch.from_numpy(v).to(dev)
    dst = torch.empty(2 * n * b, dtype=torch.float16, device=dev)
    rev = oracle.gen_rev(n, [16, 16, 2])
    perm = torch.from_numpy(rev.astype(np.int32)).to(dev)
    rc = lib.mfft_stage_convert(ctypes.c_void_p(src.data_ptr()), 3, ctypes.c_void_p(dst.data_ptr()), 0, n, b,
                                ctypes.c_void_p(perm.data_ptr()), n, None, 0, 1.0, None)
    assert rc == 0
    torch.cuda.synchronize()
    got = dst.double().cpu().numpy().reshape(b, n, 2)
    want = oracle.round_fp16(v).reshape(b, n, 2)[:, rev]
    assert np.array_equal(np.nan_to_num(got, posinf=1e9, neginf=-1e9), np.nan_to_num(want, posinf=1e9, neginf=-1e9))


@pytest.mark.gpu
def test_compat_cpp_program():
    """A C++ caller of the reference-named API, built against include/ and the library."""
    exe = os.path.join(ROOT, "fft_demo_amd", "lib", "compat_main")
    assert os.path.exists(exe), "build the library first"
    out = subprocess.run([exe, "512", "4", "32", "64"], check=True, capture_output=True, text=True, timeout=120).stdout
    print(out)
    for line in out.strip().splitlines():
        fields = dict(kv.split("=") for kv in line.split()[1:] if "=" in kv)
        assert fields["err"] == "0"
        assert float(fields["rel_l2"]) <= 4e-3, line


@pytest.mark.gpu
def test_conv_cpp_program_reference_interface():
    """fft_conv2d_half from include/tensorcore_fft_conv2d.h in a C++ program (main.cpp's
    accuracy_conv shape: 256x256 signal, 16 filters of 13x13, 'valid' correlation)."""
    exe = os.path.join(ROOT, "fft_demo_amd", "lib", "conv_main")
    assert os.path.exists(exe), "build the library first"
    r = subprocess.run([exe, "256", "13", "16"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp16", "split_fp16"])
@pytest.mark.parametrize("n,batch", [(256, 64), (512, 32), (1024, 16), (4096, 8), (65536, 4), (131072, 2),
                                     (262144, 2), (524288, 1)])
def test_digit_reversed_plans_on_fast_kernels(dev, prec, n, batch):
    """MFFT_ORDER_DIGIT_REVERSED plans (tcfftExec's input order, gen_rev of the reference's
    radix list) run on the matrix-core kernels, which read the reversed order in their
    first-stage loads: one pass up to 4096 points, two contiguous-line passes above (the
    first with its four-step twiddle on the digit-reversed line index; with many 65536-point
    lines the one-pass on-chip kernel, tests/test_gpu_onchip.py). Unnormalised, as
    tcfftExec; tolerance fp16 rel-L2 4e-3 / max 2e-2, split fp16 1e-3 / 4e-3."""
    import fft_demo_amd as F

    tol = {"fp16": (4e-3, 2e-2), "split_fp16": (1e-3, 4e-3)}[prec]
    plan = F.Plan((n,), batch, precision=prec, in_order=1, scale=0)
    kinds = [(p["kernel"], p["R"]) for p in plan.info["passes"]]
    print(f"n={n} {prec}: passes {kinds}")
    assert all(k in ("fast_rows", "fast_rows_tr") for k, _ in kinds), kinds
    assert len(kinds) == (1 if n <= 4096 else 2)
    rev = oracle.gen_rev(n, REF_RADICES[n])
    assert np.array_equal(plan.input_permutation(), rev)
    x = complex_input((batch, n), 200 + n % 89)
    y = from_device(plan.exec(to_device(x[:, rev], plan.in_dtype(), dev))) / plan.out_scale
    l2, mx = rel_errors(y, oracle.fft(x))
    print(f"  rel_l2={l2:.2e} rel_max={mx:.2e}")
    assert l2 <= tol[0] and mx <= tol[1]
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny", [(256, 256), (512, 256), (256, 1024), (1024, 1024), (64, 128)])
def test_digit_reversed_2d_plans_on_fast_kernels(dev, nx, ny):
    """2D digit-reversed plans (rows by rev_y, columns by rev_x, tcfft_half_2d.cu:1264-1271):
    the row pass and the strided column pass read the reversed orders themselves."""
    import fft_demo_amd as F

    b = 2
    plan = F.Plan((nx, ny), b, precision="fp16", in_order=1, scale=0)
    kinds = [(p["kernel"], p["R"]) for p in plan.info["passes"]]
    print(f"{nx}x{ny}: passes {kinds}")
    assert kinds == [("fast_rows", ny), ("fast_cols", nx)], kinds
    rx, ry = plan.input_permutation(0), plan.input_permutation(1)
    x = complex_input((b, nx, ny), 31)
    y = from_device(plan.exec(to_device(x[:, rx][:, :, ry], plan.in_dtype(), dev))) / plan.out_scale
    l2, mx = rel_errors(y, oracle.fft2d(x))
    print(f"  rel_l2={l2:.2e} rel_max={mx:.2e}")
    assert l2 <= 4e-3 and mx <= 2e-2
    plan.close()
