"""Pins the FP64 oracle (CPU only).

Each oracle function is checked against something that does not restate its own
formula: the reference's own digit-reversal tables, closed forms, a brute-force DFT,
an independent FFT library's vectors (numpy/pocketfft, scipy) committed as fixtures,
and the DFT's algebraic properties -- chosen so a dropped term, a wrong sign or index,
or transposed axes fail at least one check.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cplx(shape, seed):
    v = oracle.uniform(int(np.prod(shape)) * 2, seed).reshape(*shape, 2)
    return v[..., 0] + 1j * v[..., 1]


# ---------------------------------------------------------------- gen_rev
def test_gen_rev_matches_reference_tables():
    """tensorcore_fft_conv2d.h:72-120 holds rev_256 and rev_512 (zero-padded to 1024)."""
    t = json.load(open(os.path.join(GOLD, "rev_tables.json")))
    assert list(oracle.gen_rev(256, [16, 16])) == t["rev_256"][:256]
    assert list(oracle.gen_rev(512, [16, 16, 2])) == t["rev_512"][:512]
    assert all(v == 0 for v in t["rev_256"][256:]) and all(v == 0 for v in t["rev_512"][512:])


def test_gen_rev_closed_forms():
    i = np.arange(1024)
    assert np.array_equal(oracle.gen_rev(256, [16, 16]), 16 * (i[:256] % 16) + i[:256] // 16)
    assert np.array_equal(oracle.gen_rev(512, [16, 16, 2]), 32 * (i[:512] % 16) + 2 * ((i[:512] % 256) // 16) + i[:512] // 256)
    assert np.array_equal(oracle.gen_rev(1024, [16, 16, 4]), 64 * (i % 16) + 4 * ((i % 256) // 16) + i // 256)


def test_gen_rev_is_mixed_radix_digit_reversal():
    """Position i = d0 + r0*(d1 + r1*(d2 ...)) (radices in plan order, least significant
    first) holds the natural index with the digits in the opposite order:
    n = d_last + r_last*(d_{last-1} + ...)."""
    for n_, rad in [(512, [16, 16, 2]), (1024, [16, 16, 4]), (96, [4, 8, 3])]:
        rev = oracle.gen_rev(n_, rad)
        for i in range(n_):
            d, x = [], i
            for r in rad:
                d.append(x % r)
                x //= r
            n = 0
            for k in range(len(rad)):  # most significant of n is d[0]
                n = n * rad[k] + d[k]
            assert rev[i] == n, (n_, i)
        assert sorted(rev) == list(range(n_))


def test_gen_rev_rejects_bad_radices():
    with pytest.raises(ValueError):
        oracle.gen_rev(256, [16, 8])


# ---------------------------------------------------------------- data
def test_fp16_rounding_matches_numpy():
    x = np.concatenate([oracle.uniform(20000, 3, -70000, 70000), oracle.uniform(20000, 4, -1e-4, 1e-4),
                        np.array([65504.0, 65519.99, 65520.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -25, 0.0])])
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).astype(np.float64)
    assert np.array_equal(oracle.round_fp16(x), ref)


def test_bf16_rounding_ties_to_even():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3*2^-8 -> 1 + 2^-6
    x = np.array([1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -(1 + 2.0 ** -9), 1 + 2.0 ** -7])
    assert list(oracle.round_bf16(x)) == [1.0, 1 + 2.0 ** -6, -1.0, 1 + 2.0 ** -7]
    # bf16 is the top half of an fp32: truncating an exactly representable value is exact
    y = np.float32(oracle.uniform(1000, 5)).view(np.uint32) & np.uint32(0xFFFF0000)
    yv = y.view(np.float32).astype(np.float64)
    assert np.array_equal(oracle.round_bf16(yv), yv)


def test_splitmix64_published_vector():
    """The generator behind every synthetic input: SplitMix64 (Steele, Lea and Flood, OOPSLA
    2014; Vigna's reference splitmix64.c), whose published output for seed 1234567 begins
    6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
    16408922859458223821; the uniform values are (x >> 11) * 2^-53 of that stream."""
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert [int(v) for v in oracle.splitmix64(1234567, 5)] == want
    u = oracle.uniform(5, 1234567, 0.0, 1.0)
    assert [float(v) for v in u] == [(w >> 11) * 2.0 ** -53 for w in want]


def test_uniform_stream_range_and_seed():
    a = oracle.uniform(100000, 42)
    assert a.min() >= -1 and a.max() < 1 and abs(a.mean()) < 0.01
    assert np.array_equal(a, oracle.uniform(100000, 42)) and not np.array_equal(a, oracle.uniform(100000, 43))


# ---------------------------------------------------------------- DFT / FFT
@pytest.mark.parametrize("n", [1, 2, 3, 5, 12, 16, 17, 30, 256, 512, 1024, 2176, 4096])
def test_fft_matches_bruteforce_dft(n):
    x = cplx((3, n), n)
    for sign in (-1, 1):
        y = oracle.fft(x, sign=sign)
        d = oracle.dft(x, sign=sign)
        assert np.abs(y - d).max() <= 1e-11 * max(1.0, np.abs(d).max())


def test_dft_closed_forms():
    n = 64
    imp = np.zeros((1, n), complex)
    imp[0, 0] = 1
    assert np.allclose(oracle.dft(imp), np.ones(n))
    const = np.ones((1, n), complex)
    want = np.zeros(n)
    want[0] = n
    assert np.allclose(oracle.dft(const), want, atol=1e-12)
    k0 = 5
    tone = np.exp(2j * np.pi * k0 * np.arange(n) / n)[None]
    want = np.zeros(n, complex)
    want[k0] = n
    assert np.allclose(oracle.dft(tone), want, atol=1e-10)
    # forward sign: a tone at +k0 lands on bin k0, not n-k0
    assert abs(oracle.fft(tone)[0, k0]) > n - 1e-6


@pytest.mark.parametrize("n", [256, 65536])
def test_fft_properties(n):
    x = cplx((2, n), 21)
    y = cplx((2, n), 22)
    X, Y = oracle.fft(x), oracle.fft(y)
    # linearity
    assert np.abs(oracle.fft(2 * x - 3j * y) - (2 * X - 3j * Y)).max() <= 1e-9 * np.abs(X).max()
    # Parseval
    assert abs(np.sum(np.abs(X) ** 2) - n * np.sum(np.abs(x) ** 2)) <= 1e-10 * n * np.sum(np.abs(x) ** 2)
    # shift theorem: x[n-m] <-> X[k] W^(km)
    m = 37
    Xs = oracle.fft(np.roll(x, m, axis=-1))
    k = np.arange(n)
    assert np.abs(Xs - X * np.exp(-2j * np.pi * k * m / n)).max() <= 1e-9 * np.abs(X).max()
    # inverse round trip (unnormalised inverse)
    assert np.abs(oracle.fft(X, sign=1) / n - x).max() <= 1e-12 * n


def test_fft_against_independent_library_vectors():
    g = np.load(os.path.join(GOLD, "fft_vectors.npz"))
    for n in (16, 17, 256, 512, 1024, 2176, 4096):
        x = g[f"x_{n}"]
        assert np.abs(oracle.fft(x) - g[f"fwd_{n}"]).max() <= 1e-12 * n
        assert np.abs(oracle.fft(x, sign=1) - g[f"inv_{n}"]).max() <= 1e-12 * n


def test_fft2d_against_library_and_separability():
    g = np.load(os.path.join(GOLD, "fft_vectors.npz"))
    x = g["x2d_32x64"]
    y = oracle.fft2d(x)
    assert np.abs(y - g["fwd2d_32x64"]).max() <= 1e-11
    # separability: columns of the row-transformed data (nx != ny catches transposed axes)
    rows = oracle.fft(x)
    cols = np.swapaxes(oracle.fft(np.swapaxes(rows, -1, -2)), -1, -2)
    assert np.abs(y - cols).max() <= 1e-11


def test_fft_batch_loop_matches_fft():
    x = cplx((4, 1024), 31)
    a = np.stack([x.real, x.imag], -1).ravel()
    out = oracle.fft_batch_loop(a, 1024, 4).reshape(4, 1024, 2)
    assert np.abs(out[..., 0] + 1j * out[..., 1] - oracle.fft(x)).max() < 1e-10


# ---------------------------------------------------------------- convolution
def test_conv_same_against_scipy_fixtures():
    g = np.load(os.path.join(GOLD, "conv_vectors.npz"))
    for k in (3, 7, 31):
        y = oracle.conv2d_same(g["img"], g[f"filt_{k}"])
        assert np.abs(y - g[f"same_{k}"]).max() < 1e-11
        rows = oracle.conv2d_same(g["img"], g[f"filt_{k}"], rows=(10, 20))
        assert np.abs(rows - g[f"same_{k}"][10:20]).max() < 1e-11


def test_corr_valid_against_scipy_and_bruteforce():
    g = np.load(os.path.join(GOLD, "conv_vectors.npz"))
    for k in (3, 7):
        y = oracle.corr2d_valid(g["img"], g[f"filt_{k}"])
        assert np.abs(y - g[f"corr_valid_{k}"]).max() < 1e-11
    s = cplx((9, 11), 5).real
    f = cplx((3, 4), 6).real
    y = oracle.corr2d_valid(s, f)
    want = np.array([[np.sum(f * s[i:i + 3, j:j + 4]) for j in range(8)] for i in range(7)])
    assert np.abs(y - want).max() < 1e-12


def test_conv_same_is_flipped_correlation():
    """'same' convolution equals the 'valid' correlation of the zero-padded image with
    the 180-degree rotated filter (the reference's inverse_filter step,
    tensorcore_fft_conv2d_kernel.cu:1234-1253)."""
    img = cplx((20, 24), 7).real
    f = cplx((5, 5), 8).real
    pad = np.pad(img, 2)
    want = oracle.corr2d_valid(pad, f[::-1, ::-1])
    assert np.abs(oracle.conv2d_same(img, f) - want).max() < 1e-12


def test_conv_full_against_scipy_bruteforce_and_adjoint():
    """'full' convolution: scipy.signal.convolve2d (independent library), a brute-force
    sum, and the adjoint identity of the 'valid' correlation layer that makes it the
    input gradient: <corr_valid(s, w), d> = <s, conv_full(d, w)> and the filter gradient
    <corr_valid(s, w), d> = <w, corr_valid(s, d)>."""
    from scipy import signal

    a = cplx((7, 9), 11).real
    b = cplx((3, 4), 12).real
    y = oracle.conv2d_full(a, b)
    assert y.shape == (9, 12)
    assert np.abs(y - signal.convolve2d(a, b, mode="full")).max() < 1e-12
    want = np.zeros((9, 12))
    for i in range(3):
        for j in range(4):
            want[i:i + 7, j:j + 9] += b[i, j] * a
    assert np.abs(y - want).max() < 1e-12
    s = cplx((13, 10), 13).real
    w = cplx((4, 3), 14).real
    d = cplx((10, 8), 15).real  # shape of corr_valid(s, w)
    lhs = np.sum(oracle.corr2d_valid(s, w) * d)
    assert abs(lhs - np.sum(s * oracle.conv2d_full(d, w))) < 1e-10
    assert abs(lhs - np.sum(w * oracle.corr2d_valid(s, d))) < 1e-10


# ---------------------------------------------------------------- metrics
def test_err_capped_semantics():
    """main.cpp:118-134: sum of min(1,|(t-s)/s|) over re and im, / (n * batch)."""
    s = np.array([[1 + 1j, 2 - 2j, 0 + 1j]])
    t = np.array([[1.1 + 1j, 2 - 1j, 5 + 1j]])
    # re: 0.1, 0, |5/0| -> 1 ; im: 0, 0.5, 0
    assert abs(oracle.err_capped(t, s) - (0.1 + 0 + 1 + 0 + 0.5 + 0) / 3) < 1e-12


def test_err_l2max():
    s = np.array([3.0, 4.0])
    t = np.array([3.0, 4.5])
    l2, mx = oracle.err_l2max(t, s)
    assert abs(l2 - 0.5 / 5) < 1e-15 and abs(mx - 0.5 / 4) < 1e-15
