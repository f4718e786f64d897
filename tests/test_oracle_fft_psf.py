"""Pins the oracle's FFT, FFT-extent rule, pad maps and PSF preparation.

Checks come from what the mathematics fixes (brute-force DFT, explicit
padded-torus convolution, delta kernels, closed-form tables) and from numpy
as an independent library routine -- never from retyping the oracle.
"""
import numpy as np
import pytest

from conftest import relL2


@pytest.mark.parametrize("n", list(range(1, 17)) + [24, 30, 48, 60, 64, 80, 96, 105, 128, 384, 768, 1280])
def test_fft_brute_force_dft(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    assert np.abs(orc.fft(x) - dft).max() <= 1e-11 * max(1.0, np.abs(dft).max())
    idft = np.exp(2j * np.pi * np.outer(k, k) / n) @ x
    assert np.abs(orc.fft(x, inverse=True) - idft).max() <= 1e-11 * max(1.0, np.abs(idft).max())


@pytest.mark.parametrize("n", [256, 512, 1536, 2560, 3584])
def test_fft_against_numpy_and_roundtrip(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = orc.fft(x)
    assert relL2(X, np.fft.fft(x)) < 1e-14
    assert relL2(orc.fft(X, inverse=True) / n, x) < 1e-14
    # FP32 twin: accurate to single precision
    assert relL2(orc.fft(x, precision="f32"), np.fft.fft(x)) < 2e-6


def test_fft2_against_numpy(orc):
    rng = np.random.default_rng(7)
    x = rng.standard_normal((48, 80)) + 1j * rng.standard_normal((48, 80))
    assert relL2(orc.fft2(x), np.fft.fft2(x)) < 1e-14
    assert relL2(orc.fft2(x, inverse=True), np.fft.ifft2(x) * x.size) < 1e-14


# fmvd_deconvolve.cpp:28-51; values for the configurations of the workloads
SNAP = {16: 16, 17: 32, 30: 32, 33: 64, 158: 256, 286: 512, 552: 1024, 1024: 1024, 1025: 1536,
        1064: 1536, 1536: 1536, 1537: 2048, 2088: 2560, 3000: 3072, 3500: 3584, 4096: 4096}


def test_snap_transform_size_table(orc):
    for n, want in SNAP.items():
        assert orc.snap_transform_size(n) == want, n


def test_pad_index_closed_form(orc):
    n, half = 7, 3
    got = [orc.pad_index(i, n, half) for i in range(16)]
    assert got == [0, 1, 2, 3, 4, 5, 6, 6, 6, 6, 0, 0, 0, 0, 0, 0]


def _torus_conv_numpy(dp, kernel):
    """Circular convolution with the kernel centred at (0,0), via numpy FFT."""
    H, W = dp.shape
    kh, kw = kernel.shape
    kp = np.zeros((H, W))
    for i in range(kh):
        for j in range(kw):
            kp[(i - kh // 2) % H, (j - kw // 2) % W] = kernel[i, j]
    return np.real(np.fft.ifft2(np.fft.fft2(dp) * np.fft.fft2(kp)))


@pytest.mark.parametrize("shape,kshape", [((7, 9), (7, 9)), ((5, 5), (3, 5)), ((12, 6), (5, 3))])
def test_convolve_single_plane(orc, shape, kshape):
    rng = np.random.default_rng(3)
    data = rng.standard_normal(shape)
    kernel = rng.standard_normal(kshape)
    dh, dw = shape
    kh, kw = kshape
    H, W = orc.snap_transform_size(dh + kh - 1), orc.snap_transform_size(dw + kw - 1)
    rows = [orc.pad_index(y, dh, kh // 2) for y in range(H)]
    cols = [orc.pad_index(x, dw, kw // 2) for x in range(W)]
    dp = data[np.ix_(rows, cols)]
    want = _torus_conv_numpy(dp, kernel)[:dh, :dw]
    assert np.abs(orc.convolve_single_plane(data, kernel) - want).max() < 1e-12


def test_convolve_delta_is_identity(orc):
    rng = np.random.default_rng(4)
    data = rng.standard_normal((9, 11))
    delta = np.zeros((5, 7))
    delta[2, 3] = 1.0
    assert np.array_equal(orc.convolve_single_plane(data, delta), data)


def _kernels(V=3, kh=7, kw=5, seed=9):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.01, 1.0, size=(V, kh, kw)).astype(np.float32)


def test_kernel_hat_type0_is_flip(orc):
    k = _kernels()
    kh = orc.kernel_hat(k, 0)
    for v in range(k.shape[0]):
        f = k[v, ::-1, ::-1].astype(np.float64)
        assert np.allclose(kh[v], f / f.sum(), rtol=1e-14, atol=0)


def test_kernel_hat_type3_is_flipped_power(orc):
    k = _kernels()
    V = k.shape[0]
    kh = orc.kernel_hat(k, 3)
    for v in range(V):
        f = (k[v].astype(np.float64) ** V)[::-1, ::-1]
        assert np.allclose(kh[v], f / f.sum(), rtol=1e-13, atol=0)


@pytest.mark.parametrize("psf", [1, 2])
def test_kernel_hat_compound_against_numpy(orc, psf):
    k = _kernels(V=3, kh=5, kw=7)
    got = orc.kernel_hat(k, psf)
    kd = k.astype(np.float64)
    V, kh, kw = k.shape
    H, W = orc.snap_transform_size(2 * kh - 1), orc.snap_transform_size(2 * kw - 1)
    rows = [orc.pad_index(y, kh, kh // 2) for y in range(H)]
    cols = [orc.pad_index(x, kw, kw // 2) for x in range(W)]

    def conv(a, b):
        return _torus_conv_numpy(a[np.ix_(rows, cols)], b)[:kh, :kw]

    for v in range(V):
        k2 = kd[v, ::-1, ::-1].copy()
        for w in range(V):
            if w == v:
                continue
            t = conv(kd[v, ::-1, ::-1], kd[w])
            if psf == 1:
                t = conv(t, kd[w, ::-1, ::-1])
            k2 *= t
        assert np.allclose(got[v], k2 / k2.sum(), rtol=1e-11, atol=1e-15)


def test_normalize_f32_semantics(orc):
    k = _kernels(V=1)[0]
    got = orc.normalize_f32(k)
    s = np.float32(np.sum(k.astype(np.float64)))
    assert np.array_equal(got, (k / s).astype(np.float32))
    assert abs(float(got.astype(np.float64).sum()) - 1.0) < 1e-6


def test_normalized_weights_float_semantics(orc):
    rng = np.random.default_rng(11)
    w = rng.uniform(0, 1, size=(3, 6, 5)).astype(np.float32)
    w[:, 0, 0] = 0.0
    got = orc.normalized_weights(w)
    s = np.float32(0)
    for v in range(3):
        s = (s + w[v]).astype(np.float32)
    want = np.where(s > 0, w / np.where(s > 0, s, 1), w).astype(np.float32)
    assert np.array_equal(got, want)
    assert np.all(got[:, 0, 0] == 0)
