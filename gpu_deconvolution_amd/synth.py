"""Deterministic synthetic multi-view SPIM data (beads, Gaussian PSFs, Poisson).

Used by the tests and the benchmark.  Everything is seeded: seed = 1000 +
workload index by convention, one numpy PCG64 stream for the bead layout and
one per view for the noise.  Layout is [z][y][x].

Ground truth: background 10 plus Gaussian beads (sigma U(1,3), peak amplitude
U(100,1000)) rendered within +-ceil(4 sigma).  View v is the ground truth
blurred by a separable 3D Gaussian whose (sx, sy, sz) is the base PSF rotated
v*90 degrees about y (x and z swap on odd views), then Poisson noise, clamped
to the sample range.  The per-plane RL kernel of a view is the normalised xy
marginal of its 3D PSF on a k x k grid.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Workload:
    shape: tuple          # (Z, H, W)
    views: int
    beads: int
    kernel: int           # kernel edge length
    sigma: tuple = (1.5, 1.5, 4.0)  # base PSF (sx, sy, sz)
    seed: int = 1000
    bits: int = 16


def view_sigma(base, v):
    sx, sy, sz = base
    return (sz, sy, sx) if v % 2 else (sx, sy, sz)


def gaussian_kernel_2d(sx: float, sy: float, k: int) -> np.ndarray:
    r = np.arange(k, dtype=np.float64) - k // 2
    gx = np.exp(-0.5 * (r / sx) ** 2)
    gy = np.exp(-0.5 * (r / sy) ** 2)
    ker = np.outer(gy, gx)
    return (ker / ker.sum()).astype(np.float32)


def render_beads(shape, n, rng, amp=(100.0, 1000.0), sig=(1.0, 3.0), background=10.0) -> np.ndarray:
    Z, H, W = shape
    vol = np.full(shape, background, np.float32)
    cz = rng.uniform(0, Z, n)
    cy = rng.uniform(0, H, n)
    cx = rng.uniform(0, W, n)
    s = rng.uniform(sig[0], sig[1], n)
    a = rng.uniform(amp[0], amp[1], n)
    for i in range(n):
        r = int(math.ceil(4 * s[i]))
        z0, z1 = max(0, int(cz[i]) - r), min(Z, int(cz[i]) + r + 1)
        y0, y1 = max(0, int(cy[i]) - r), min(H, int(cy[i]) + r + 1)
        x0, x1 = max(0, int(cx[i]) - r), min(W, int(cx[i]) + r + 1)
        if z0 >= z1 or y0 >= y1 or x0 >= x1:
            continue
        zz = (np.arange(z0, z1) - cz[i])[:, None, None]
        yy = (np.arange(y0, y1) - cy[i])[None, :, None]
        xx = (np.arange(x0, x1) - cx[i])[None, None, :]
        vol[z0:z1, y0:y1, x0:x1] += (a[i] * np.exp(-(zz * zz + yy * yy + xx * xx) / (2 * s[i] ** 2))).astype(np.float32)
    return vol


def blur_separable(vol: np.ndarray, sigma) -> np.ndarray:
    """Separable Gaussian blur with reflect-free zero-flux 'nearest' borders."""
    from scipy.ndimage import gaussian_filter1d
    sx, sy, sz = sigma
    out = gaussian_filter1d(vol, sx, axis=2, mode="nearest", truncate=4.0)
    out = gaussian_filter1d(out, sy, axis=1, mode="nearest", truncate=4.0)
    out = gaussian_filter1d(out, sz, axis=0, mode="nearest", truncate=4.0)
    return out


def make(w: Workload):
    """Returns dict(truth, views (V,Z,H,W) uint, kernels (V,k,k) f32, weights (V,H,W) f32)."""
    rng = np.random.default_rng(w.seed)
    truth = render_beads(w.shape, w.beads, rng)
    maxv = (1 << w.bits) - 1
    views, kernels = [], []
    for v in range(w.views):
        s = view_sigma(w.sigma, v)
        blurred = blur_separable(truth, s)
        nrng = np.random.default_rng([w.seed, v])
        noisy = nrng.poisson(np.maximum(blurred, 0).astype(np.float64))
        views.append(np.clip(noisy, 0, maxv).astype(np.uint16 if w.bits == 16 else np.uint8))
        kernels.append(gaussian_kernel_2d(s[0], s[1], w.kernel))
    Z, H, W = w.shape
    weights = np.ones((w.views, H, W), np.float32)
    return dict(truth=truth, views=np.stack(views), kernels=np.stack(kernels), weights=weights)


def planes_of(views: np.ndarray):
    """(V, Z, H, W) -> list over z of per-view planes (the datasource order)."""
    V, Z = views.shape[:2]
    return [[views[v, z] for v in range(V)] for z in range(Z)]


def random_affine(rng, shape, max_rot_deg=15.0, scale=(0.9, 1.1), max_shift=20.0) -> np.ndarray:
    """Row-major 3x4 affine about the volume centre (rotations about x, y, z,
    isotropic scale, translation), float32."""
    Z, H, W = shape
    ax, ay, az = np.deg2rad(rng.uniform(-max_rot_deg, max_rot_deg, 3))
    s = rng.uniform(*scale)
    t = rng.uniform(-max_shift, max_shift, 3)
    Rx = np.array([[1, 0, 0], [0, math.cos(ax), -math.sin(ax)], [0, math.sin(ax), math.cos(ax)]])
    Ry = np.array([[math.cos(ay), 0, math.sin(ay)], [0, 1, 0], [-math.sin(ay), 0, math.cos(ay)]])
    Rz = np.array([[math.cos(az), -math.sin(az), 0], [math.sin(az), math.cos(az), 0], [0, 0, 1]])
    A = s * (Rz @ Ry @ Rx)
    c = np.array([(W - 1) / 2, (H - 1) / 2, (Z - 1) / 2])
    b = c - A @ c + t
    return np.concatenate([A, b[:, None]], 1).astype(np.float32)


def invert_affine(m: np.ndarray) -> np.ndarray:
    M = np.eye(4)
    M[:3] = np.asarray(m, np.float64).reshape(3, 4)
    return np.linalg.inv(M)[:3].astype(np.float32)


# ----------------------------------------------------------------- GPU side
_synth = None


def synth_lib():
    """libfmvd_synth.so: GPU generator for large workloads (same HIP runtime
    as the reconstruction library)."""
    global _synth
    if _synth is None:
        import ctypes as C
        import os
        from . import lib
        lib()
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfmvd_synth.so")
        L = C.CDLL(path)
        vp, i = C.c_void_p, C.c_int
        L.fmvd_synth_truth.argtypes = [vp, i, i, i, vp, i, C.c_float, vp]
        L.fmvd_synth_blur.argtypes = [vp, vp, i, i, i, i, C.c_float, vp]
        L.fmvd_synth_poisson.argtypes = [vp, vp, C.c_long, C.c_ulonglong, i, vp]
        L.fmvd_synth_copy.argtypes = [vp, vp, C.c_long, vp]
        _synth = L
    return _synth


def gpu_views(shape, views, beads, seed, sigma=(1.5, 1.5, 4.0), bits=16):
    """Ground truth with `beads` beads, blurred per view (base PSF rotated
    v*90 degrees about y) and Poisson-noised, on the device.  Returns a list
    of DeviceArrays (Z, H, W) of samples, one per view."""
    import ctypes as C
    from . import DeviceArray, check, device_synchronize
    L = synth_lib()
    Z, H, W = shape
    rng = np.random.default_rng(seed)
    b = np.stack([rng.uniform(0, Z, beads), rng.uniform(0, H, beads), rng.uniform(0, W, beads),
                  rng.uniform(1.0, 3.0, beads), rng.uniform(100.0, 1000.0, beads)], 1).astype(np.float32)
    truth = DeviceArray(shape, np.float32)
    t1 = DeviceArray(shape, np.float32)
    t2 = DeviceArray(shape, np.float32)
    L.fmvd_synth_truth(C.c_void_p(truth.ptr), Z, H, W, b.ctypes.data_as(C.c_void_p), beads, 10.0, None)
    check()
    out = []
    dt = np.uint16 if bits == 16 else np.uint8
    for v in range(views):
        sx, sy, sz = view_sigma(sigma, v)
        L.fmvd_synth_blur(C.c_void_p(truth.ptr), C.c_void_p(t1.ptr), Z, H, W, 0, sx, None)
        L.fmvd_synth_blur(C.c_void_p(t1.ptr), C.c_void_p(t2.ptr), Z, H, W, 1, sy, None)
        L.fmvd_synth_blur(C.c_void_p(t2.ptr), C.c_void_p(t1.ptr), Z, H, W, 2, sz, None)
        d = DeviceArray(shape, dt)
        L.fmvd_synth_poisson(C.c_void_p(t1.ptr), C.c_void_p(d.ptr), Z * H * W, (seed << 8) + v, bits, None)
        check()
        out.append(d)
    device_synchronize()
    for x in (truth, t1, t2):
        x.free()
    return out
