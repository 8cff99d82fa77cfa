"""GPU parity of the conv header's FFT entry points (include/tensorcore_fft_conv2d.h) called
from a C++ program built against it (fft_demo_amd/lib/conv_api_main), against the FP64
oracle:

  * TensorCoreFFTCreate / TensorCoreFFTExec (tensorcore_fft_conv2d_kernel.cu:1495-1648):
    forward (fft = 1, stream overload) and inverse (fft = -1, plain overload), input
    digit-reversed along both axes with gen_rev of the reference's radix lists (its
    rev_256 / rev_512 tables, tensorcore_fft_conv2d.h:72-120), fp16 matrix-core,
    unnormalised; tolerance rel-L2 4e-3 / rel-max 2e-2 (fp16 data);
  * fft_2d_c2c / ifft_2d_c2c (tensorcore_fft_conv2d.cpp:79-214) with batch % num_streams
    != 0 (the reference drops that remainder; here every transform is done), two calls on
    one object; FP64 host data, split fp16 with fp32 device buffers, unnormalised;
    tolerance rel-L2 1e-4 / rel-max 1e-3;
  * the 10-argument conv2d_forward (tensorcore_fft_conv2d.cpp:291-363) on a
    default-constructed object: 'valid' correlation of S channels with F x S filters
    (main.cpp:97-116); tolerance rel-L2 1e-4 per output map.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.helpers import complex_input

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fft_demo_amd", "lib", "conv_api_main")
REF_LISTS = {64: [16, 4], 128: [16, 8], 256: [16, 16], 512: [16, 16, 2], 1024: [16, 16, 4]}


def _run(args, tmp_path):
    assert os.path.exists(EXE), "build the library first"
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr


def _interleave(x):
    return np.ascontiguousarray(np.stack([x.real, x.imag], axis=-1), dtype=np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny,batch,fft", [(256, 256, 2, 1), (512, 512, 1, -1), (256, 512, 2, -1), (512, 256, 1, 1),
                                             (128, 64, 3, 1)])
def test_tensorcore_fft_create_exec(tmp_path, nx, ny, batch, fft):
    x = complex_input((batch, nx, ny), 400 + nx + ny + fft)  # fp16-exact values
    rx, ry = oracle.gen_rev(nx, REF_LISTS[nx]), oracle.gen_rev(ny, REF_LISTS[ny])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _interleave(x[:, rx][:, :, ry]).tofile(fin)
    _run(["tcfft", nx, ny, batch, fft, fin, fout], tmp_path)
    y = np.fromfile(fout, dtype=np.float64).reshape(batch, nx, ny, 2)
    y = y[..., 0] + 1j * y[..., 1]
    ref = oracle.fft2d(x, sign=-1 if fft == 1 else 1)
    l2, mx = oracle.err_l2max(y, ref)
    print(f"TensorCoreFFT {nx}x{ny} b={batch} fft={fft}: rel_l2={l2:.2e} rel_max={mx:.2e}")
    assert l2 <= 4e-3 and mx <= 2e-2


@pytest.mark.gpu
@pytest.mark.parametrize("x,y,batch,streams,fft", [(128, 64, 5, 3, 1), (64, 96, 7, 4, -1), (256, 256, 3, 8, 1)])
def test_fft_2d_c2c_uneven_streams(tmp_path, x, y, batch, streams, fft):
    v = oracle.uniform(batch * y * x * 2, 500 + x + batch, -1.0, 1.0)
    data = v.reshape(batch, y, x, 2)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np.ascontiguousarray(data).tofile(fin)
    _run(["c2c", x, y, batch, streams, fft, fin, fout], tmp_path)
    out = np.fromfile(fout, dtype=np.float64).reshape(batch, y, x, 2)
    got = out[..., 0] + 1j * out[..., 1]
    ref = oracle.fft2d(data[..., 0] + 1j * data[..., 1], sign=-1 if fft == 1 else 1)
    for b in range(batch):  # the last block is a short one
        l2, mx = oracle.err_l2max(got[b], ref[b])
        assert l2 <= 1e-4 and mx <= 1e-3, (b, l2, mx)
    l2, mx = oracle.err_l2max(got, ref)
    print(f"fft_2d_c2c x={x} y={y} b={batch} streams={streams} fft={fft}: rel_l2={l2:.2e} rel_max={mx:.2e}")


@pytest.mark.gpu
@pytest.mark.parametrize("S,F,sy,sx,fy,fx", [(2, 3, 60, 50, 7, 5), (1, 16, 256, 256, 13, 13)])
def test_conv2d_forward_long_overload(tmp_path, S, F, sy, sx, fy, fx):
    sig = oracle.uniform(S * sy * sx, 600 + S, 0.0, 1.0).reshape(S, sy, sx)
    filt = oracle.uniform(F * S * fy * fx, 700 + F, -1.0, 1.0).reshape(F, S, fy, fx)
    fs, ff, fo = tmp_path / "s.bin", tmp_path / "f.bin", tmp_path / "o.bin"
    sig.tofile(fs)
    filt.tofile(ff)
    _run(["conv10", S, F, sy, sx, fy, fx, fs, ff, fo], tmp_path)
    oh, ow = sy - fy + 1, sx - fx + 1
    out = np.fromfile(fo, dtype=np.float64).reshape(F, oh, ow)
    for f in range(F):
        ref = sum(oracle.corr2d_valid(sig[s], filt[f, s]) for s in range(S))
        l2, mx = oracle.err_l2max(out[f], ref)
        assert l2 <= 1e-4 and mx <= 1e-3, (f, l2, mx)
    print(f"conv2d_forward (10 args) S={S} F={F} {sy}x{sx} filter {fy}x{fx}: ok")
