"""Writes the golden fixtures in tests/golden/*.npz.

An independent numpy restatement of the reference's plane-wise multi-view RL
(bene51/gpu_deconvolution), used to pin the C oracle:
  * FFT extents: fmvd_deconvolve.cpp:22-51
  * pad maps: fmvd_deconvolve_cuda.cu:96-117 (weights), fmvd_deconvolve_cuda.impl.cu:39-59
    (data, half-widths swapped via fmvd_deconvolve.impl.cpp:36-37)
  * kernel padding: fmvd_deconvolve_cuda.cu:27-37
  * PSF types: fmvd_deconvolve.cpp:139-204, with convolve_single_plane (:90-137) done here
    by numpy FFTs on the clamp-padded torus (the C oracle sums directly)
  * weights: fmvd_deconvolve.impl.cpp:86-124 (float32 sum in view order, divide where > 0)
  * RL iteration: fmvd_deconvolve.impl.cpp:262-273, ratio guard Q > 1
    (fmvd_deconvolve_cuda.impl.cu:173-175), update order fmvd_deconvolve_cuda.cu:387-391
  * unpad: fmvd_deconvolve_cuda.impl.cu:121-126
Every convolution here goes through numpy.fft (a library routine), so the
fixtures check the oracle's own FFT and direct sums against an independent path.

Run:  python tests/golden/make_golden.py
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def snap(n):
    n = n if n % 16 == 0 else n - n % 16 + 16
    p = 1 << (n.bit_length() - 1)
    if p == n:
        return n
    if 2 * p <= 1024:
        return 2 * p
    return n if n % 512 == 0 else n - n % 512 + 512


def pad_map(n_pad, n, half):
    i = np.arange(n_pad)
    return np.where(i < n, i, np.where(i < n + half, n - 1, 0))


def pad_kernel(k, H, W):
    kh, kw = k.shape
    out = np.zeros((H, W), np.float64)
    for i in range(kh):
        for j in range(kw):
            out[(i - kh // 2) % H, (j - kw // 2) % W] = k[i, j]
    return out


def torus_conv(x, kp):
    return np.real(np.fft.ifft2(np.fft.fft2(x) * np.fft.fft2(kp)))


def convolve_single_plane(data, kernel):
    dh, dw = data.shape
    kh, kw = kernel.shape
    H, W = snap(dh + kh - 1), snap(dw + kw - 1)
    padded = data[np.ix_(pad_map(H, dh, kh // 2), pad_map(W, dw, kw // 2))]
    return torus_conv(padded, pad_kernel(kernel, H, W))[:dh, :dw]


def flip(k):
    return k[::-1, ::-1]


def kernel_hat(raw, psf_type):
    raw = raw.astype(np.float64)
    V = raw.shape[0]
    out = []
    for v in range(V):
        if psf_type == 0:
            k2 = flip(raw[v]).copy()
        elif psf_type in (1, 2):
            k2 = flip(raw[v]).copy()
            for w in range(V):
                if w == v:
                    continue
                t1 = convolve_single_plane(flip(raw[v]), raw[w])
                if psf_type == 1:
                    t1 = convolve_single_plane(t1, flip(raw[w]))
                k2 = k2 * t1
        else:
            k2 = flip(raw[v] ** V)
        out.append(k2 / k2.sum())
    return np.stack(out)


def normalize_f32(k):
    s = np.sum(k.astype(np.float64))
    return (k / np.float32(s)).astype(np.float32)


def norm_weights(w):
    s = np.zeros(w.shape[1:], np.float32)
    for v in range(w.shape[0]):
        s = (s + w[v]).astype(np.float32)
    out = np.where(s > 0, w / np.where(s > 0, s, 1), w).astype(np.float32)
    return out


def rl(data, weights, kernels, psf_type, iterations, bits):
    Z, V, dh, dw = data.shape
    kh, kw = kernels.shape[1:]
    H, W = snap(dh + kh - 1), snap(dw + kw - 1)
    khat = kernel_hat(kernels, psf_type)
    kn = np.stack([normalize_f32(kernels[v]) for v in range(V)]).astype(np.float64)
    wn = norm_weights(weights).astype(np.float64)
    ry, rx = pad_map(H, dh, kh // 2), pad_map(W, dw, kw // 2)
    dy, dx = pad_map(H, dh, kw // 2), pad_map(W, dw, kh // 2)
    Wp = np.stack([wn[v][np.ix_(ry, rx)] for v in range(V)])
    K = np.stack([np.fft.fft2(pad_kernel(kn[v], H, W)) for v in range(V)])
    Kh = np.stack([np.fft.fft2(pad_kernel(khat[v], H, W)) for v in range(V)])
    ests, outs = [], []
    maxv = (1 << bits) - 1
    for z in range(Z):
        Dp = np.stack([data[z, v].astype(np.float64)[np.ix_(dy, dx)] for v in range(V)])
        E = np.full((H, W), 10.0 * V)
        for _ in range(iterations):
            for v in range(V):
                Q = np.real(np.fft.ifft2(np.fft.fft2(E) * K[v]))
                R = np.where(Q > 1, Dp[v] / np.where(Q > 1, Q, 1), Q)
                Cc = np.real(np.fft.ifft2(np.fft.fft2(R) * Kh[v]))
                E = E + Wp[v] * (E * Cc - E)
        e = E[:dh, :dw]
        ests.append(e)
        t = np.minimum(e, maxv)
        outs.append(np.where(t > 0, np.floor(t), 0).astype(np.uint16))
    return np.stack(outs), np.stack(ests), khat


def case(name, Z, V, dh, dw, kh, kw, psf_type, iterations, bits, seed, weight_zero=False):
    rng = np.random.default_rng(seed)
    maxv = (1 << bits) - 1
    data = rng.integers(0, min(maxv, 900), size=(Z, V, dh, dw)).astype(np.uint16)
    weights = rng.uniform(0.2, 1.0, size=(V, dh, dw)).astype(np.float32)
    if weight_zero:
        weights[:, : dh // 3, : dw // 4] = 0.0
    r = np.arange(kh)[:, None] - kh // 2
    c = np.arange(kw)[None, :] - kw // 2
    kernels = np.stack([np.exp(-0.5 * ((r / (1.2 + 0.4 * v)) ** 2 + (c / (1.5 + 0.3 * v)) ** 2))
                        + 0.01 * rng.uniform(size=(kh, kw)) for v in range(V)]).astype(np.float32)
    out, est, khat = rl(data, weights, kernels, psf_type, iterations, bits)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), data=data, weights=weights, kernels=kernels,
                        psf_type=psf_type, iterations=iterations, bits=bits, out=out, est=est, khat=khat)
    print(name, out.shape, float(est.mean()))


def transform_case(name, seed):
    """Trilinear transform fixture: expected values from exact rational arithmetic."""
    from fractions import Fraction
    rng = np.random.default_rng(seed)
    d, h, w = 5, 7, 9
    src = rng.integers(0, 65535, size=(d, h, w)).astype(np.uint16)
    src[0, 0, 0] = 65535
    src[1, 2, 3] = 32768
    m = np.array([[0.97, 0.05, -0.02, 0.75], [-0.04, 1.02, 0.03, -0.4], [0.01, -0.03, 0.99, 0.35]], np.float32)
    td, th, tw = 4, 6, 8
    out = np.zeros((td, th, tw), np.uint16)

    def T(x, y, z):
        if 0 <= x < w and 0 <= y < h and 0 <= z < d:
            return Fraction(int(src[z, y, x]))
        return Fraction(0)

    mf = [Fraction(float(v)) for v in m.reshape(-1)]
    for z in range(td):
        for y in range(th):
            for x in range(tw):
                r = [mf[4 * a] * x + mf[4 * a + 1] * y + mf[4 * a + 2] * z + mf[4 * a + 3] for a in range(3)]
                fl = [math.floor(ri) for ri in r]
                a = [ri - f for ri, f in zip(r, fl)]
                val = Fraction(0)
                for dz in (0, 1):
                    for dy in (0, 1):
                        for dx in (0, 1):
                            wgt = ((a[0] if dx else 1 - a[0]) * (a[1] if dy else 1 - a[1]) *
                                   (a[2] if dz else 1 - a[2]))
                            val += wgt * T(fl[0] + dx, fl[1] + dy, fl[2] + dz)
                yv = val * 65536 / 65535 + Fraction(1, 2)
                o = int(yv.numerator // yv.denominator)
                out[z, y, x] = min(o, 65535)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), src=src, m=m, out=out)
    print(name, out.shape)


if __name__ == "__main__":
    case("rl_psf0_16", Z=2, V=2, dh=24, dw=20, kh=5, kw=7, psf_type=0, iterations=3, bits=16, seed=1)
    case("rl_psf1_16", Z=1, V=3, dh=20, dw=26, kh=7, kw=5, psf_type=1, iterations=2, bits=16, seed=2)
    case("rl_psf2_8", Z=2, V=2, dh=18, dw=30, kh=5, kw=5, psf_type=2, iterations=3, bits=8, seed=3)
    case("rl_psf3_16", Z=1, V=2, dh=33, dw=17, kh=9, kw=3, psf_type=3, iterations=2, bits=16, seed=4)
    case("rl_wzero_16", Z=1, V=2, dh=22, dw=24, kh=5, kw=5, psf_type=0, iterations=4, bits=16, seed=5,
         weight_zero=True)
    transform_case("transform_affine_16", seed=6)
