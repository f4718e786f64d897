"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product.  See oracle.h for the
reference file:line each function restates and for what pins it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        i, vp = C.c_int, C.c_void_p
        L.orc_snap_transform_size.argtypes = [i]
        L.orc_pad_index.argtypes = [i, i, i]
        L.orc_fft_c2c_f64.argtypes = [_dp, i, i]
        L.orc_fft_c2c_f32.argtypes = [_fp, i, i]
        L.orc_fft2_c2c_f64.argtypes = [_dp, i, i, i]
        L.orc_normalize_f32.argtypes = [_fp, i]
        L.orc_normalize_f32.restype = None
        L.orc_flip_kernel_f64.argtypes = [_dp, _dp, i, i]
        L.orc_flip_kernel_f64.restype = None
        L.orc_convolve_single_plane.argtypes = [_dp, i, i, _dp, i, i]
        L.orc_convolve_single_plane.restype = None
        L.orc_kernel_hat.argtypes = [_fp, i, i, i, i, _dp]
        L.orc_normalized_weights.argtypes = [_fp, i, i, i, _fp]
        L.orc_normalized_weights.restype = None
        L.orc_rl_torus_f64.argtypes = [i, i, i, _dp, _dp, _dp, _dp, i, _dp]
        L.orc_rl_torus_f32.argtypes = [i, i, i, _fp, _fp, _fp, _fp, i, _fp]
        for n in ("orc_rl_planes_f64", "orc_rl_planes_f32"):
            getattr(L, n).argtypes = [i, i, i, i, i, _u16p, _fp, _fp, i, i, i, i, vp, vp, vp, i]
        L.orc_build_operands_f64.argtypes = [i, i, i, vp, _fp, _fp, i, i, i, i, i, _dp, _dp, _dp, _dp]
        L.orc_transform.argtypes = [_u16p, i, i, i, i, _fp, i, i, i, i, _u16p]
        L.orc_transform.restype = None
        L.orc_mask.argtypes = [i, i, i, i, i, i, _fp, i, C.c_float, _fp]
        L.orc_mask.restype = None
        _lib = L
    return _lib


def snap_transform_size(n: int) -> int:
    return lib().orc_snap_transform_size(int(n))


def pad_index(i: int, n: int, half: int) -> int:
    return lib().orc_pad_index(int(i), int(n), int(half))


def fft(x: np.ndarray, inverse: bool = False, precision: str = "f64") -> np.ndarray:
    """Unnormalised complex FFT of a 1-D array."""
    n = x.shape[0]
    if precision == "f64":
        buf = np.ascontiguousarray(np.stack([x.real, x.imag], -1).reshape(-1), dtype=np.float64)
        rc = lib().orc_fft_c2c_f64(buf, n, int(inverse))
    else:
        buf = np.ascontiguousarray(np.stack([x.real, x.imag], -1).reshape(-1), dtype=np.float32)
        rc = lib().orc_fft_c2c_f32(buf, n, int(inverse))
    assert rc == 0
    return buf[0::2] + 1j * buf[1::2]


def fft2(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    h, w = x.shape
    buf = np.ascontiguousarray(np.stack([x.real, x.imag], -1).reshape(-1), dtype=np.float64)
    assert lib().orc_fft2_c2c_f64(buf, h, w, int(inverse)) == 0
    return (buf[0::2] + 1j * buf[1::2]).reshape(h, w)


def normalize_f32(k: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(k, dtype=np.float32).copy()
    lib().orc_normalize_f32(k.reshape(-1), k.size)
    return k


def convolve_single_plane(data: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    d = np.ascontiguousarray(data, dtype=np.float64).copy()
    k = np.ascontiguousarray(kernel, dtype=np.float64)
    dh, dw = d.shape
    kh, kw = k.shape
    lib().orc_convolve_single_plane(d.reshape(-1), dw, dh, k.reshape(-1), kw, kh)
    return d


def kernel_hat(kernels: np.ndarray, psf_type: int) -> np.ndarray:
    """kernels: (V, kh, kw) raw float32 -> (V, kh, kw) float64 normalised kernel-hat."""
    k = np.ascontiguousarray(kernels, dtype=np.float32)
    V, kh, kw = k.shape
    out = np.zeros((V, kh, kw), np.float64)
    assert lib().orc_kernel_hat(k.reshape(-1), kw, kh, V, int(psf_type), out.reshape(-1)) == 0
    return out


def normalized_weights(weights: np.ndarray) -> np.ndarray:
    w = np.ascontiguousarray(weights, dtype=np.float32)
    V, dh, dw = w.shape
    out = np.zeros_like(w)
    lib().orc_normalized_weights(w.reshape(-1), dw, dh, V, out.reshape(-1))
    return out


def build_operands(plane_views, weights, kernels, psf_type, ffth, fftw):
    """Padded Dp, Wp, kp, kph (each (V, ffth, fftw) float64) as the reference builds them.
    kernels (V, kh, kw) float32 are normalised in place."""
    w = np.ascontiguousarray(weights, dtype=np.float32)
    V, dh, dw = w.shape
    kh, kw = kernels.shape[1:]
    outs = [np.zeros((V, ffth, fftw), np.float64) for _ in range(4)]
    pv = None
    if plane_views is not None:
        pv = np.ascontiguousarray(plane_views, dtype=np.uint16)
    rc = lib().orc_build_operands_f64(dw, dh, V, pv.ctypes.data if pv is not None else None, w.reshape(-1),
                                      kernels.reshape(-1), kw, kh, int(psf_type), ffth, fftw,
                                      *[o.reshape(-1) for o in outs])
    assert rc == 0
    return outs


def rl_torus(Dp, Wp, kp, kph, iterations, E0, precision="f64"):
    V, h, w = Dp.shape
    if precision == "f64":
        E = np.ascontiguousarray(E0, dtype=np.float64).copy()
        args = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (Dp, Wp, kp, kph)]
        rc = lib().orc_rl_torus_f64(h, w, V, *args, int(iterations), E.reshape(-1))
    else:
        E = np.ascontiguousarray(E0, dtype=np.float32).copy()
        args = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in (Dp, Wp, kp, kph)]
        rc = lib().orc_rl_torus_f32(h, w, V, *args, int(iterations), E.reshape(-1))
    assert rc == 0
    return E


def rl_planes(data, weights, kernels, psf_type, iterations, bits=16, init=None, precision="f64", threads=1):
    """data: (Z, V, dh, dw) samples; weights (V, dh, dw); kernels (V, kh, kw) float32 raw
    (a normalised copy is used).  Returns (samples (Z,dh,dw) uint16, estimate (Z,dh,dw) f64)."""
    d = np.ascontiguousarray(data, dtype=np.uint16)
    Z, V, dh, dw = d.shape
    k = np.ascontiguousarray(kernels, dtype=np.float32).copy()
    kh, kw = k.shape[1:]
    w = np.ascontiguousarray(weights, dtype=np.float32)
    est = np.zeros((Z, dh, dw), np.float64)
    out = np.zeros((Z, dh, dw), np.uint16)
    ini = None if init is None else np.ascontiguousarray(init, dtype=np.float64)
    fn = lib().orc_rl_planes_f64 if precision == "f64" else lib().orc_rl_planes_f32
    rc = fn(dw, dh, Z, V, bits, d.reshape(-1), w.reshape(-1), k.reshape(-1), kw, kh, int(psf_type),
            int(iterations), ini.ctypes.data if ini is not None else None, est.ctypes.data, out.ctypes.data,
            int(threads))
    assert rc == 0
    return out, est


def transform(src: np.ndarray, inv: np.ndarray, target_shape, bits=16, z0=0, nz=None) -> np.ndarray:
    s = np.ascontiguousarray(src, dtype=np.uint16)
    d, h, w = s.shape
    td, th, tw = target_shape
    nz = td - z0 if nz is None else nz
    out = np.zeros((nz, th, tw), np.uint16)
    m = np.ascontiguousarray(inv, dtype=np.float32).reshape(12)
    lib().orc_transform(s.reshape(-1), w, h, d, bits, m, tw, th, z0, nz, out.reshape(-1))
    return out


def mask(src_shape, target_shape, inv, border, zspacing) -> np.ndarray:
    d, h, w = src_shape
    td, th, tw = target_shape
    out = np.zeros((th, tw), np.float32)
    m = np.ascontiguousarray(inv, dtype=np.float32).reshape(12)
    lib().orc_mask(w, h, d, tw, th, td, m, int(border), float(zspacing), out.reshape(-1))
    return out
