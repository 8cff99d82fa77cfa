"""GPU test of mfft_exec_noalloc (include/mfft.h), the exec entry point that allocates
nothing: the plan's intermediates go to the library's per-GPU scratch array instead of a
workspace. Called through the C ABI, its output is compared with the existing entry
point's (mfft_exec for the whole batch, mfft_exec_ex for part of it) on the same inputs:
bit for bit (the same kernels on the same data), and against the FP64 oracle at the split
fp16 tolerance (rel-L2 <= 1e-3, rel-max <= 4e-3). Two calls in a row, the second on a
smaller input; a plan whose intermediates exceed the array is refused; the calls leave the
device's free memory unchanged (no device allocation).
"""
import ctypes

import pytest

import oracle
from tests.helpers import complex_input, from_device, rel_errors, to_device

MFFT_ERR_NOMEM = -4
SCRATCH_BYTES = 64 << 20


def _noalloc(lib, plan, x, out, batch, stream):
    return lib.mfft_exec_noalloc(plan._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), int(batch),
                                 ctypes.c_void_p(stream.cuda_stream))


@pytest.mark.gpu
@pytest.mark.parametrize("shape,batch,small,mid,split", [((65536,), 32, 5, "c16", 256), ((65536,), 256, 5, "c16", 0),
                                                         ((1024, 1024), 8, 3, None, 0), ((4096,), 64, 17, None, 0)])
def test_exec_noalloc_matches_exec(dev, shape, batch, small, mid, split):
    """split = 256 at 65536 points: the two-pass plan (intermediates in the scratch array);
    256 lines with split = 0: the one-pass on-chip plan, which has no intermediates in memory,
    and whose exec of 5 lines runs its companion two-pass plan (intermediates in the array)."""
    import torch

    import fft_demo_amd as F
    from fft_demo_amd import _lib as L

    lib = L.load()
    # the plan keeps its intermediates within the scratch array (a 32 MiB chunk)
    plan = F.Plan(shape, batch, precision="split_fp16", mid_fmt=mid, workspace_limit=32 << 20, split_hint=split)
    assert plan.info["workspace_bytes"] <= SCRATCH_BYTES
    s = torch.cuda.current_stream(dev)
    x = complex_input((batch, *shape), 71)
    xd = to_device(x, plan.in_dtype(), dev)
    want = plan.exec(xd)  # mfft_exec, the plan's own workspace
    x2 = complex_input((small, *shape), 72)
    xd2 = torch.zeros_like(xd)
    xd2[:small] = to_device(x2, plan.in_dtype(), dev)
    want2 = torch.zeros_like(xd)
    L.check(lib.mfft_exec_ex(plan._h, ctypes.c_void_p(xd2.data_ptr()), ctypes.c_void_p(want2.data_ptr()), small, None,
                             ctypes.c_void_p(s.cuda_stream)), "mfft_exec_ex")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(dev)[0]
    got = torch.full_like(want, 7.0)
    got2 = torch.zeros_like(want)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info(dev)[0]
    assert _noalloc(lib, plan, xd, got, batch, s) == 0, lib.mfft_last_error()
    assert _noalloc(lib, plan, xd2, got2, small, s) == 0, lib.mfft_last_error()  # second call, smaller input
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info(dev)[0]
    print(f"{shape} x {batch}: workspace {plan.info['workspace_bytes']} B, free memory "
          f"{free0} -> {free1} -> {free2}")
    assert free2 == free1, "mfft_exec_noalloc changed the device's free memory"
    assert torch.equal(got, want)
    assert torch.equal(got2[:small], want2[:small])
    assert float(got2[small:].abs().max()) == 0.0  # transforms past the requested batch untouched
    for xs, ys in ((x, got), (x2, got2[:small])):
        y = from_device(ys) / plan.out_scale
        ref = oracle.fft(xs) if len(shape) == 1 else oracle.fft2d(xs)
        l2, mx = rel_errors(y, ref)
        print(f"  rel_l2={l2:.2e} rel_max={mx:.2e}")
        assert l2 <= 1e-3 and mx <= 4e-3
    plan.close()


@pytest.mark.gpu
def test_exec_noalloc_refuses_plans_beyond_the_scratch(dev):
    """65536 x 256 with fp32 intermediates in one chunk needs 128 MiB: refused with a clear
    error; the same shape with a 64 MiB workspace limit runs."""
    import torch

    import fft_demo_amd as F
    from fft_demo_amd import _lib as L

    lib = L.load()
    big = F.Plan((65536,), 256, precision="split_fp16", workspace_limit=128 << 20)
    assert big.info["workspace_bytes"] > SCRATCH_BYTES
    xd = torch.zeros((256, 65536, 2), dtype=torch.float16, device=dev)
    out = torch.empty_like(xd)
    s = torch.cuda.current_stream(dev)
    assert _noalloc(lib, big, xd, out, 256, s) == MFFT_ERR_NOMEM
    assert "scratch" in lib.mfft_last_error().decode()
    big.close()
    ok = F.Plan((65536,), 256, precision="split_fp16", workspace_limit=SCRATCH_BYTES)
    assert ok.info["workspace_bytes"] <= SCRATCH_BYTES
    x = complex_input((256, 65536), 73)
    xd = to_device(x, torch.float16, dev)
    want = ok.exec(xd)
    assert _noalloc(lib, ok, xd, out, 256, s) == 0, lib.mfft_last_error()
    torch.cuda.synchronize()
    assert torch.equal(out, want)
    ok.close()
