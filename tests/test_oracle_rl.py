"""Pins the oracle's RL iteration (orc_rl_torus_*, orc_rl_planes_*).

Known-answer tests whose answers follow from the mathematics of RL on the
torus (uniform data, fixed points, flux conservation, zero weights,
transpose symmetry), a numpy-FFT re-implementation, and the golden fixtures
written by tests/golden/make_golden.py.
"""
import glob
import os

import numpy as np
import pytest

from conftest import relL2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _pad_kernel(k, H, W):
    kh, kw = k.shape
    out = np.zeros((H, W))
    for i in range(kh):
        for j in range(kw):
            out[(i - kh // 2) % H, (j - kw // 2) % W] = k[i, j]
    return out


def _gauss(kh, kw, sy, sx):
    r = np.arange(kh)[:, None] - kh // 2
    c = np.arange(kw)[None, :] - kw // 2
    g = np.exp(-0.5 * ((r / sy) ** 2 + (c / sx) ** 2))
    return g / g.sum()


def test_single_view_uniform_data_one_iteration(orc):
    # KAT (i): V=1, w=1, D=c>1 everywhere: E0=10 -> Q=10 -> R=c/10 -> C=c/10 -> E=c
    H, W, c = 32, 48, 137.0
    k = _gauss(5, 7, 1.3, 2.0)
    kp = _pad_kernel(k, H, W)[None]
    kph = _pad_kernel(k[::-1, ::-1], H, W)[None]
    E = orc.rl_torus(np.full((1, H, W), c), np.ones((1, H, W)), kp, kph, 1, np.full((H, W), 10.0))
    assert np.abs(E - c).max() < 1e-10 * c


def test_uniform_data_scalar_recurrence(orc):
    # KAT (ii): uniform data c_v per view and uniform weights w_v: every plane
    # stays uniform and follows E <- E + w_v (c_v - E) while E > 1.
    H, W, V = 16, 32, 3
    cs = [50.0, 200.0, 120.0]
    ws = [0.2, 0.5, 0.3]
    ks = [_gauss(5, 5, 1.0 + 0.5 * v, 1.5) for v in range(V)]
    kp = np.stack([_pad_kernel(k, H, W) for k in ks])
    kph = np.stack([_pad_kernel(k[::-1, ::-1], H, W) for k in ks])
    Dp = np.stack([np.full((H, W), c) for c in cs])
    Wp = np.stack([np.full((H, W), w) for w in ws])
    E = orc.rl_torus(Dp, Wp, kp, kph, 4, np.full((H, W), 10.0 * V))
    e = 10.0 * V
    for _ in range(4):
        for v in range(V):
            e = e + ws[v] * (cs[v] - e)
    assert np.abs(E - e).max() < 1e-9 * e


def test_fixed_point(orc):
    # KAT (iii): D = x (*) k on the torus, V=1, w=1, khat = flip(k): x is invariant.
    rng = np.random.default_rng(1)
    H, W = 32, 32
    x = rng.uniform(50, 500, size=(H, W))
    k = _gauss(7, 5, 1.5, 1.1)
    kp = _pad_kernel(k, H, W)
    D = np.real(np.fft.ifft2(np.fft.fft2(x) * np.fft.fft2(kp)))
    E = orc.rl_torus(D[None], np.ones((1, H, W)), kp[None], _pad_kernel(k[::-1, ::-1], H, W)[None], 3, x)
    assert relL2(E, x) < 1e-12


def test_flux_conservation(orc):
    # KAT (iv): V=1, w=1, Q>1 everywhere, normalised k, khat=flip(k): sum E1 = sum D
    rng = np.random.default_rng(2)
    H, W = 32, 64
    D = rng.uniform(20, 900, size=(H, W))
    k = _gauss(9, 9, 2.0, 1.2)
    E = orc.rl_torus(D[None], np.ones((1, H, W)), _pad_kernel(k, H, W)[None],
                     _pad_kernel(k[::-1, ::-1], H, W)[None], 1, np.full((H, W), 10.0))
    assert abs(E.sum() - D.sum()) < 1e-10 * D.sum()


def test_torus_matches_numpy_fft_restatement(orc):
    rng = np.random.default_rng(3)
    H, W, V = 48, 40, 2
    Dp = rng.uniform(0, 700, size=(V, H, W))
    Wp = rng.uniform(0, 1, size=(V, H, W))
    Wp /= Wp.sum(0)
    ks = [_gauss(5, 7, 1.2, 2.1), _gauss(7, 5, 2.0, 1.0)]
    kp = np.stack([_pad_kernel(k, H, W) for k in ks])
    kph = np.stack([_pad_kernel(k[::-1, ::-1], H, W) for k in ks])
    E0 = np.full((H, W), 20.0)
    got = orc.rl_torus(Dp, Wp, kp, kph, 3, E0)
    E = E0.copy()
    for _ in range(3):
        for v in range(V):
            Q = np.real(np.fft.ifft2(np.fft.fft2(E) * np.fft.fft2(kp[v])))
            R = np.where(Q > 1, Dp[v] / np.where(Q > 1, Q, 1), Q)
            C = np.real(np.fft.ifft2(np.fft.fft2(R) * np.fft.fft2(kph[v])))
            E = E + Wp[v] * (E * C - E)
    assert relL2(got, E) < 1e-12
    # the FP32 twin agrees to single-precision accuracy
    got32 = orc.rl_torus(Dp, Wp, kp, kph, 3, E0, precision="f32")
    assert relL2(got32, E) < 1e-5


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "rl_*.npz"))))
def test_rl_planes_golden(orc, path):
    g = np.load(path)
    kernels = g["kernels"].copy()
    out, est = orc.rl_planes(g["data"], g["weights"], kernels, int(g["psf_type"]), int(g["iterations"]),
                             bits=int(g["bits"]))
    assert relL2(est, g["est"]) < 1e-12
    assert np.array_equal(out, g["out"])
    assert np.allclose(orc.kernel_hat(g["kernels"], int(g["psf_type"])), g["khat"], rtol=1e-12, atol=1e-16)


def test_zero_weight_pixels_keep_initial_estimate(orc):
    # KAT (v): where every view's weight is 0 the update is 0 -> E stays 10*V
    g = np.load(os.path.join(GOLDEN, "rl_wzero_16.npz"))
    _, est = orc.rl_planes(g["data"], g["weights"], g["kernels"].copy(), 0, 4)
    zero = np.all(g["weights"] == 0, axis=0)
    assert zero.any()
    assert np.all(est[0][zero] == 20.0)


def test_transpose_symmetry(orc):
    # KAT (vi): transposing data, weights and kernels transposes the result
    rng = np.random.default_rng(5)
    data = rng.integers(0, 800, size=(1, 2, 20, 28)).astype(np.uint16)
    w = rng.uniform(0.1, 1, size=(2, 20, 28)).astype(np.float32)
    k = np.stack([_gauss(5, 7, 1.1, 1.9), _gauss(5, 7, 2.2, 1.3)]).astype(np.float32)
    _, e1 = orc.rl_planes(data, w, k.copy(), 2, 3)
    _, e2 = orc.rl_planes(data.transpose(0, 1, 3, 2), w.transpose(0, 2, 1),
                          k.transpose(0, 2, 1).copy(), 2, 3)
    assert relL2(e2[0].T, e1[0]) < 1e-12


def test_rl_planes_threads_and_f32_twin(orc):
    g = np.load(os.path.join(GOLDEN, "rl_psf0_16.npz"))
    _, e1 = orc.rl_planes(g["data"], g["weights"], g["kernels"].copy(), 0, 3, threads=1)
    _, e2 = orc.rl_planes(g["data"], g["weights"], g["kernels"].copy(), 0, 3, threads=2)
    assert np.array_equal(e1, e2)
    _, e32 = orc.rl_planes(g["data"], g["weights"], g["kernels"].copy(), 0, 3, precision="f32")
    assert relL2(e32, e1) < 1e-5


def test_rl_planes_kernels_normalised_in_place_like_reference(orc):
    # fmvd_deconvolve.impl.cpp:63-64 normalises the caller's kernels
    import ctypes
    g = np.load(os.path.join(GOLDEN, "rl_psf0_16.npz"))
    k = np.ascontiguousarray(g["kernels"], dtype=np.float32).copy()
    before = k.copy()
    d = np.ascontiguousarray(g["data"], dtype=np.uint16)
    Z, V, dh, dw = d.shape
    est = np.zeros((Z, dh, dw))
    orc.lib().orc_rl_planes_f64(dw, dh, Z, V, 16, d.reshape(-1), g["weights"].reshape(-1), k.reshape(-1),
                                k.shape[2], k.shape[1], 0, 1, None, est.ctypes.data, None, 1)
    for v in range(V):
        s = np.float32(np.sum(before[v].astype(np.float64)))
        assert np.array_equal(k[v], (before[v] / s).astype(np.float32))
