"""Host-side logic of the request server, no GPU: the workload generator and the C ABI's
argument checks (which reject before any device call)."""
import ctypes

import numpy as np
import pytest

from fft_demo_amd import serve


def test_trace_mix_and_ranges():
    t = serve.make_trace(rate=2000.0, duration=5.0, seed=42)
    assert 9000 < len(t) < 11000  # Poisson(10000)
    kinds = np.array([r[1] for r in t])
    frac = [np.mean(kinds == k) for k in (serve.KIND_1D, serve.KIND_2D, serve.KIND_CONV)]
    assert abs(frac[0] - 0.60) < 0.02 and abs(frac[1] - 0.25) < 0.02 and abs(frac[2] - 0.15) < 0.02
    times = np.array([r[0] for r in t])
    assert np.all(np.diff(times) > 0) and times[-1] < 5.0
    gaps = np.diff(times)
    assert abs(gaps.mean() - 1 / 2000) < 0.05 / 2000 * 10
    for _, kind, n, nx, ny, b, fid in t:
        if kind == serve.KIND_1D:
            assert n & (n - 1) == 0 and 256 <= n <= 1 << 20 and 1 <= b <= 64
        elif kind == serve.KIND_2D:
            assert nx == ny and nx in serve.SIDES_2D and 1 <= b <= 16
        else:
            assert nx == ny and nx in serve.SIDES_CONV and 0 <= fid < serve.N_FILTERS
    assert serve.make_trace(2000.0, 1.0, 42) == serve.make_trace(2000.0, 1.0, 42)


def test_filters():
    fs = serve.make_filters()
    assert len(fs) == serve.N_FILTERS
    sizes = sorted({f.shape[0] for f in fs})
    assert sizes[0] == 3 and sizes[-1] == 63 and all(s % 2 == 1 for s in sizes)
    for f in fs:
        assert f.shape[0] == f.shape[1]
        assert abs(f.sum() - 1.0) < 0.01  # normalised, then +-1% noise
        c = f.shape[0] // 2
        assert f[c, c] == pytest.approx(f.max(), rel=0.03)


def test_server_rejects_bad_descriptors_without_gpu():
    from fft_demo_amd import _lib

    lib = _lib.load()
    for kw in ({"streams": 0}, {"streams": 17}, {"hbm_budget": 0}, {"max_batch_1d": 0}):
        with pytest.raises(serve.ServerError):
            serve.Server(**kw)
    r = _lib.Request()
    assert lib.mfft_server_submit(None, ctypes.byref(r)) == -1
    assert lib.mfft_server_poll(None, None, 0, None) == -1
    assert lib.mfft_server_destroy(None) == 0
