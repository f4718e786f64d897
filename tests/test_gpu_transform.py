"""GPU parity of the affine transform and mask (through the C ABI): every
output sample bit-exact against the FP64-coordinate oracle, for both bit
depths, through the GPU-resident entry point and through transform_cuda_*'s
file output; the mask plane within one float ulp of the FP64 oracle."""
import os

import numpy as np
import pytest

from gpu_util import empty_device, from_device, to_device

pytestmark = pytest.mark.gpu

IDENT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)


def affine(seed, shape, rot=15.0):
    from gpu_deconvolution_amd.synth import random_affine
    return random_affine(np.random.default_rng(seed), shape, max_rot_deg=rot, max_shift=4.0)


def volume(shape, bits, seed):
    rng = np.random.default_rng(seed)
    maxv = (1 << bits) - 1
    v = rng.integers(0, maxv + 1, size=shape)
    v.reshape(-1)[:8] = [0, maxv, maxv - 1, 1 << (bits - 1), (1 << (bits - 1)) - 1, 1, 2, 3]
    return v.astype(np.uint16 if bits == 16 else np.uint8)


MATS = {
    "identity": lambda s: IDENT,
    "shift": lambda s: np.array([[1, 0, 0, 3], [0, 1, 0, -2], [0, 0, 1, 1]], np.float32),
    "halfvoxel": lambda s: np.array([[1, 0, 0, 0.5], [0, 1, 0, 0.5], [0, 0, 1, -0.5]], np.float32),
    "affine1": lambda s: affine(1, s),
    "affine2": lambda s: affine(2, s),
    "scale": lambda s: np.array([[1.1, 0, 0, -1.0], [0, 0.93, 0, 0.7], [0, 0, 1.05, 0.2]], np.float32),
}


@pytest.mark.parametrize("bits", [16, 8])
@pytest.mark.parametrize("name", list(MATS))
def test_transform_device_bit_exact(fm, orc, bits, name):
    shape = (24, 40, 56)
    tshape = (20, 44, 50)
    src = volume(shape, bits, seed=len(name) + bits)
    M = MATS[name](shape)
    print(f"\n[transform] bits={bits} mat={name}", flush=True)
    dt = np.uint16 if bits == 16 else np.uint8
    dsrc = to_device(src)
    ddst = empty_device(tshape, dt)
    fm.transform_device(dsrc.ptr, shape, ddst.ptr, tshape, M, bits=bits,
                        stream=0)
    fm.device_synchronize()
    got = from_device(ddst, dt)
    want = orc.transform(src.astype(np.uint16), M, tshape, bits=bits).astype(dt)
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))


def test_transform_file_api_and_mask(fm, orc, tmp_path):
    shape = (30, 33, 47)
    tshape = (26, 35, 41)
    src = volume(shape, 16, seed=5)
    M = affine(7, shape)
    out = str(tmp_path / "out.raw")
    maskf = str(tmp_path / "mask.raw")
    fm.transform_to_file(src, M, tshape, out, mask_file=maskf, border=6, zspacing=2.5)
    got = np.fromfile(out, np.uint16).reshape(tshape)
    want = orc.transform(src, M, tshape)
    assert np.array_equal(got, want)
    mask = np.fromfile(maskf, np.float32).reshape(tshape[1:])
    mwant = orc.mask(shape, tshape, M, 6, 2.5)
    assert np.abs(mask - mwant).max() <= 2 * np.finfo(np.float32).eps
    assert (mask > 0).any() and (mask == 0).any()


def test_transform_file_api_8bit_many_chunks(fm, orc, tmp_path):
    # enough target planes that the driver writes several double-buffered chunks
    shape = (20, 300, 310)
    tshape = (2100, 256, 256)
    src = volume(shape, 8, seed=8)
    M = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0.01, 0]], np.float32)
    out = str(tmp_path / "o8.raw")
    fm.transform_to_file(src, M, tshape, out)
    got = np.fromfile(out, np.uint8).reshape(tshape)
    for z in (0, 1, 1023, 1024, 2099):
        want = orc.transform(src.astype(np.uint16), M, tshape, bits=8, z0=z, nz=1)[0].astype(np.uint8)
        assert np.array_equal(got[z], want), z


def test_mask_device(fm, orc):
    shape, tshape = (40, 50, 60), (30, 52, 64)
    M = affine(3, shape)
    d = empty_device(tshape[1:], np.float32)
    fm.mask_device(shape, tshape, M, 10, 1.5, d.ptr, stream=0)
    fm.device_synchronize()
    got = from_device(d, np.float32)
    assert np.abs(got - orc.mask(shape, tshape, M, 10, 1.5)).max() <= 2 * np.finfo(np.float32).eps


def test_transform_global_gather_fallback(fm, orc):
    """A map that magnifies by 2 and rotates by 45 deg about z makes every
    16 x 16 x 8 tile's source box exceed the largest LDS brick, so the kernel
    gathers from global memory; results stay bit-exact."""
    shape = (40, 120, 130)
    tshape = (18, 40, 44)
    src = volume(shape, 16, seed=12)
    c, s = np.cos(np.pi / 4), np.sin(np.pi / 4)
    M = np.array([[2 * c, -2 * s, 0.0, 60.0], [2 * s, 2 * c, 0.0, 10.0], [0.0, 0.0, 2.0, 1.5]], np.float32)
    dsrc = to_device(src)
    ddst = empty_device(tshape, np.uint16)
    fm.transform_device(dsrc.ptr, shape, ddst.ptr, tshape, M, bits=16, stream=0)
    fm.device_synchronize()
    got = from_device(ddst, np.uint16)
    want = orc.transform(src, M, tshape, bits=16)
    assert (want > 0).mean() > 0.5
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))
