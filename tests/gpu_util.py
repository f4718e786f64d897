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
