"""Learned warm start for the multi-view RL: training data, training loop and
inference, all on the GPU through the library's C ABI (fmvd_net.h,
fmvd_batch.h, fmvd_deconvolve.h).

Training pairs are (views -> plain-RL estimate) that the plain method itself
generates: bead volumes blurred per view and Poisson-noised (synth.gpu_views),
labelled by `label_iters` iterations of the batched RL.  The network sees the
views divided by a per-sample scale (their mean) and learns the label over the
same scale.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import Batch, DeviceArray, FmvdError, Net, check, device_synchronize, lib, synth


class Sample:
    def __init__(self, views, kernels, weights, scale):
        self.views = views            # list of DeviceArray (D, H, W) uint16
        self.kernels = kernels        # list of (k, k) float32 (raw)
        self.weights = weights        # list of (H, W) float32
        self.scale = float(scale)
        self.target = None            # DeviceArray (D, H, W) float32, label / scale

    def free(self):
        for v in self.views:
            v.free()
        if self.target is not None:
            self.target.free()


def make_samples(n, shape=(64, 128, 128), nviews=2, kernel=31, label_iters=50, seed=0, group=48,
                 beads_per_mvox=500, log=None):
    """n synthetic samples with RL labels.  Bead count per volume follows the
    small workload's density (500 beads in 256 x 256 x 64)."""
    D, H, W = shape
    beads = max(10, int(round(beads_per_mvox * D * H * W / (256 * 256 * 64))))
    rng = np.random.default_rng(seed)
    samples = []
    t0 = time.time()
    for i in range(n):
        sxy, sz = float(rng.uniform(1.3, 1.7)), float(rng.uniform(3.5, 4.5))
        views = synth.gpu_views(shape, nviews, beads=beads, seed=seed * 100003 + i, sigma=(sxy, sxy, sz))
        kernels = [synth.gaussian_kernel_2d(*synth.view_sigma((sxy, sxy, sz), v)[:2], kernel)
                   for v in range(nviews)]
        weights = [np.ones((H, W), np.float32) for _ in range(nviews)]
        scale = float(np.mean([v.download().mean(dtype=np.float64) for v in views]))
        samples.append(Sample(views, kernels, weights, max(scale, 1.0)))
    for g0 in range(0, n, group):
        part = samples[g0:g0 + group]
        vols = []
        for s in part:
            est = DeviceArray(shape, np.float32)
            vols.append(dict(views=s.views, weights=s.weights, kernels=[k.copy() for k in s.kernels], out=None,
                             est=est))
        with Batch(vols) as b:
            b.run(label_iters, use_graph=True)
            device_synchronize()
        for s, v in zip(part, vols):
            host = v["est"].download() / np.float32(s.scale)
            s.target = DeviceArray.from_host(host.astype(np.float32))
            v["est"].free()
    if log:
        log(f"[warmstart] {n} samples of {shape} x {nviews} views, labels = {label_iters} RL iterations, "
            f"{time.time() - t0:.1f}s")
    return samples


class Trainer:
    """Mini-batch Adam on the GPU: the batch's network input is assembled by
    fmvd_net_prep_input from the samples' view volumes, targets are copied
    into one contiguous buffer."""

    def __init__(self, net: Net, shape, nviews, batch=8):
        self.net, self.shape, self.nviews, self.batch = net, tuple(shape), nviews, batch
        D, H, W = shape
        self.x = DeviceArray((batch, D, H, W, 2), np.uint16)
        self.t = DeviceArray((batch, D, H, W), np.float32)
        self.ptrs = DeviceArray((batch * nviews,), np.uint64)
        self.inv = DeviceArray((batch,), np.float32)
        L = lib()
        L.fmvd_net_prep_input.restype = C.c_int
        L.fmvd_net_prep_input.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_void_p]

    def step(self, samples, lr):
        D, H, W = self.shape
        n = len(samples)
        assert n == self.batch
        self.ptrs.upload(np.array([v.ptr for s in samples for v in s.views], np.uint64))
        self.inv.upload(np.array([1.0 / s.scale for s in samples], np.float32))
        if lib().fmvd_net_prep_input(self.ptrs.ptr, self.nviews, n, D, H, W, self.inv.ptr, self.x.ptr, None) != 0:
            check()
        per = D * H * W * 4
        for i, s in enumerate(samples):
            lib().fmvd_memcpy(C.c_void_p(self.t.ptr + i * per), C.c_void_p(s.target.ptr), per, 3)
        check()
        loss = self.net.backward(self.x, self.t, (n, D, H, W))
        self.net.adam(lr)
        return loss

    def fit(self, samples, epochs, lr=1e-3, seed=0, log=None, lr_decay=0.5):
        rng = np.random.default_rng(seed)
        hist = []
        t0 = time.time()
        for ep in range(epochs):
            perm = rng.permutation(len(samples))
            losses = []
            for b in range(len(samples) // self.batch):
                idx = perm[b * self.batch:(b + 1) * self.batch]
                losses.append(self.step([samples[i] for i in idx], lr))
            hist.append(float(np.mean(losses)))
            if log:
                log(f"[warmstart] epoch {ep}: mean loss {hist[-1]:.4e} lr {lr:.2e} ({time.time() - t0:.1f}s)")
            lr *= lr_decay
        return hist

    def free(self):
        for b in (self.x, self.t, self.ptrs, self.inv):
            b.free()


def warm_start(net: Net, views, scale, out: DeviceArray, floor_value=None):
    """Network estimate of a volume (tile-wise inference), in data units."""
    if floor_value is None:
        floor_value = 1e-3 * scale
    net.infer(views, scale, out, floor_value=floor_value)
    return out


def view_scale(views) -> float:
    """Per-volume scale: the mean sample value over all views."""
    return max(1.0, float(np.mean([v.download().mean(dtype=np.float64) for v in views])))


__all__ = ["Sample", "make_samples", "Trainer", "warm_start", "view_scale", "FmvdError"]
