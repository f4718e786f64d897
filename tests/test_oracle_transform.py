"""Pins the oracle's affine transform and mask (orc_transform, orc_mask).

Checks: the identity and its p -> p+1 quirk for p >= 2^(B-1) (the
2^B/(2^B-1) scaling of fmvd_transform.impl.cu:52-53), integer translations
(a shift with zero fill), half-voxel translation (exact neighbour means),
the exact-rational trilinear fixture from tests/golden, and the mask's closed
form at chosen points.
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
IDENT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)


def quirk(p, bits=16):
    """floor(p * 2^B/(2^B-1) + 1/2) clamped: exact integer arithmetic."""
    m = (1 << bits) - 1
    return np.minimum((2 * p * (m + 1) + m) // (2 * m), m)


@pytest.mark.parametrize("bits", [8, 16])
def test_identity_quirk(orc, bits):
    m = (1 << bits) - 1
    vals = np.arange(m + 1, dtype=np.int64)
    n = vals.size
    side = int(math.ceil(math.sqrt(n)))
    src = np.zeros((1, side, side), np.uint16)
    src.reshape(-1)[:n] = vals
    out = orc.transform(src, IDENT, src.shape, bits=bits)
    want = quirk(src.astype(np.int64), bits)
    assert np.array_equal(out.astype(np.int64), want)
    half = 1 << (bits - 1)
    flat = out.reshape(-1)[:n].astype(np.int64)
    assert np.array_equal(flat[:half], vals[:half])
    assert np.array_equal(flat[half:m], vals[half:m] + 1)
    assert flat[m] == m


def test_integer_translation(orc):
    rng = np.random.default_rng(1)
    src = rng.integers(0, 30000, size=(6, 7, 9)).astype(np.uint16)
    M = IDENT.copy()
    M[:, 3] = [2, -1, 1]  # target (x,y,z) samples source (x+2, y-1, z+1)
    out = orc.transform(src, M, src.shape).astype(np.int64)
    want = np.zeros_like(out)
    d, h, w = src.shape
    for z in range(d):
        for y in range(h):
            for x in range(w):
                sx, sy, sz = x + 2, y - 1, z + 1
                if 0 <= sx < w and 0 <= sy < h and 0 <= sz < d:
                    want[z, y, x] = int(src[sz, sy, sx])
    assert np.array_equal(out, quirk(want))


def test_half_voxel_translation_is_exact_mean(orc):
    rng = np.random.default_rng(2)
    src = rng.integers(0, 60000, size=(3, 4, 8)).astype(np.uint16)
    M = IDENT.copy()
    M[0, 3] = 0.5
    out = orc.transform(src, M, src.shape).astype(np.int64)
    s = src.astype(np.int64)
    right = np.concatenate([s[:, :, 1:], np.zeros_like(s[:, :, :1])], axis=2)
    tot = s + right  # interp = tot / 2 exactly
    # floor(tot/2 * 65536/65535 + 1/2) in exact integers
    want = np.minimum((tot * 65536 + 65535) // (2 * 65535), 65535)
    assert np.array_equal(out, want)


def test_exact_rational_fixture(orc):
    g = np.load(os.path.join(GOLDEN, "transform_affine_16.npz"))
    out = orc.transform(g["src"], g["m"], g["out"].shape)
    assert np.array_equal(out, g["out"])


def test_transform_z_subset(orc):
    rng = np.random.default_rng(3)
    src = rng.integers(0, 60000, size=(8, 10, 12)).astype(np.uint16)
    M = np.array([[0.9, 0.1, 0.05, 1.3], [-0.1, 0.95, 0.02, 0.7], [0.03, -0.04, 1.05, -0.6]], np.float32)
    full = orc.transform(src, M, (8, 10, 12))
    part = orc.transform(src, M, (8, 10, 12), z0=3, nz=4)
    assert np.array_equal(full[3:7], part)


def test_mask_closed_form(orc):
    w, h, d, border = 40, 30, 20, 8
    M = IDENT.copy()
    out = orc.mask((d, h, w), (d, h, w), M, border, 2.0)
    z = d // 2
    dz = min(z, d - z) * 2.0

    def f(t):
        return 0.5 * (1 - math.cos(t / border * math.pi)) if t < border else 1.0

    for (y, x) in [(0, 0), (3, 5), (15, 20), (29, 39), (10, 1)]:
        dx = x if x < w // 2 else w - x
        dy = y if y < h // 2 else h - y
        want = f(dx) * f(dy) * f(dz)
        assert abs(float(out[y, x]) - want) < 1e-7
    M2 = IDENT.copy()
    M2[0, 3] = -5  # x < 5 maps outside the source
    out2 = orc.mask((d, h, w), (d, h, w), M2, border, 2.0)
    assert np.all(out2[:, :5] == 0)
