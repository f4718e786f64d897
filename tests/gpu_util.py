"""Helpers for the GPU parity tests: device buffers from the library's own
HIP runtime (fmvd_device.h) and the comparison rules against the oracle."""
import numpy as np

from conftest import relL2

# FP32 estimate vs the FP64 oracle: relative L2 over the data region.
RL_TOL = 1e-4
# Truncated samples may differ by one LSB only where the oracle estimate lies
# this close to an integer boundary.
BOUNDARY = 1e-2


def to_device(arr: np.ndarray):
    import gpu_deconvolution_amd as fm
    return fm.DeviceArray.from_host(arr)


def from_device(d, dtype=None):
    return d.download()


def empty_device(shape, dtype):
    import gpu_deconvolution_amd as fm
    return fm.DeviceArray(shape, dtype)


def check_rl(out_gpu, est_gpu, out_orc, est_orc, tol=RL_TOL, bits=16):
    """Returns (relL2, number of one-LSB sample differences at boundaries)."""
    r = relL2(est_gpu, est_orc)
    assert r <= tol, f"relL2 {r:.3e} > {tol:.1e}"
    diff = np.abs(out_gpu.astype(np.int64) - out_orc.astype(np.int64))
    assert diff.max() <= 1, f"sample difference {diff.max()}"
    frac = est_orc - np.floor(est_orc)
    near = (frac < BOUNDARY) | (frac > 1 - BOUNDARY) | (est_orc < 1) | (est_orc > (1 << bits) - 2)
    bad = int(np.count_nonzero((diff > 0) & ~near))
    assert bad == 0, f"{bad} sample mismatches away from integer boundaries"
    return r, int(np.count_nonzero(diff))


def make_case(Z, V, dh, dw, kh, kw, bits, seed, beads=40):
    """Seeded bead-field planes (Z, V, dh, dw) with Poisson noise, anisotropic
    Gaussian kernels (V, kh, kw) and uniform weights (V, dh, dw)."""
    rng = np.random.default_rng(seed)
    maxv = (1 << bits) - 1
    yy, xx = np.mgrid[0:dh, 0:dw]
    base = np.full((dh, dw), 10.0)
    for _ in range(beads):
        cy, cx = rng.uniform(0, dh), rng.uniform(0, dw)
        s = rng.uniform(1.5, 4.0)
        base += rng.uniform(100, 900) * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
    data = np.empty((Z, V, dh, dw), np.uint16)
    for z in range(Z):
        for v in range(V):
            lam = base * (0.6 + 0.1 * v + 0.05 * z)
            if bits == 8:
                lam = lam * 0.25
            data[z, v] = np.clip(rng.poisson(lam), 0, maxv)
    r = np.arange(kh)[:, None] - kh // 2
    c = np.arange(kw)[None, :] - kw // 2
    kernels = np.stack([np.exp(-0.5 * ((r / (1.3 + 0.6 * v)) ** 2 + (c / (2.0 - 0.3 * v)) ** 2))
                        for v in range(V)]).astype(np.float32)
    weights = rng.uniform(0.3, 1.0, size=(V, dh, dw)).astype(np.float32)
    return data, kernels, weights
