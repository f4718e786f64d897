"""Python mirror of the C ABI of libNativeSPIMReconstructionCuda.so (gfx950).

The product is the shared library built from ``csrc/``; this module only
loads it with ctypes and mirrors the reference interface
(bene51/gpu_deconvolution ``fmvd_deconvolve.impl.h:6-66``,
``fmvd_transform.h:4-24``, ``fmvd_cuda_utils.h:4-25``) so tests and the
benchmark call exactly what a C or JNI caller calls.  There is no CPU
fallback: importing ``lib()`` raises if the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Callable, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libNativeSPIMReconstructionCuda.so")

PSF_TYPES = {"independent": 0, "efficient_bayesian": 1, "optimization_1": 2, "optimization_2": 3}

ERROR_HANDLER_T = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p)
SOURCE16_T = C.CFUNCTYPE(C.c_int, C.POINTER(C.POINTER(C.c_uint16)), C.c_int)
SINK16_T = C.CFUNCTYPE(None, C.POINTER(C.c_uint16))
SOURCE8_T = C.CFUNCTYPE(C.c_int, C.POINTER(C.POINTER(C.c_uint8)), C.c_int)
SINK8_T = C.CFUNCTYPE(None, C.POINTER(C.c_uint8))
ESINK_T = C.CFUNCTYPE(None, C.POINTER(C.c_float), C.c_int, C.c_void_p)

TIMING_KINDS = {"column": 0, "ratio": 1, "update": 2, "ingest": 3, "egress": 4, "transform": 5}

_lib = None
_errors: list = []


class FmvdError(RuntimeError):
    pass


def _record_error(_param, msg):
    _errors.append(msg.decode() if msg else "")


_error_cb = ERROR_HANDLER_T(_record_error)


def lib() -> C.CDLL:
    """Load the library (raises if it was not built) and install a
    non-exiting error handler that records messages for :func:`check`."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                          "(make -C gpu_deconvolution_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp, i, f, cp = C.c_void_p, C.c_int, C.c_float, C.c_char_p
    fpp = C.POINTER(C.POINTER(C.c_float))
    sigs = {
        "setErrorHandler": (None, [ERROR_HANDLER_T, vp]),
        "iDivUp": (i, [i, i]),
        "getNumCUDADevices": (i, []),
        "getCudaDeviceName": (None, [i, cp]),
        "setCudaDevice": (None, [i]),
        "fmvd_last_error": (i, [cp, i]),
        "fmvd_malloc": (vp, [C.c_size_t]),
        "fmvd_free": (None, [vp]),
        "fmvd_initialize_cuda_16": (vp, [i, i, fpp, fpp, i, i, i, i, i, SOURCE16_T, SINK16_T]),
        "fmvd_initialize_cuda_8": (vp, [i, i, fpp, fpp, i, i, i, i, i, SOURCE8_T, SINK8_T]),
        "fmvd_deconvolve_planes_cuda_16": (None, [vp, i]),
        "fmvd_deconvolve_planes_cuda_8": (None, [vp, i]),
        "fmvd_destroy_cuda_16": (None, [vp]),
        "fmvd_destroy_cuda_8": (None, [vp]),
        "fmvd_deconvolve_files_cuda_16": (None, [vp, vp, i, i, i, fpp, fpp, i, i, i, i, i]),
        "fmvd_deconvolve_files_cuda_8": (None, [vp, vp, i, i, i, fpp, fpp, i, i, i, i, i]),
        "fmvd_set_estimate_sink_cuda_16": (i, [vp, ESINK_T, vp]),
        "fmvd_set_estimate_sink_cuda_8": (i, [vp, ESINK_T, vp]),
        "fmvd_set_batch_planes_cuda_16": (i, [vp, i]),
        "fmvd_set_batch_planes_cuda_8": (i, [vp, i]),
        "fmvd_set_chunk_planes_cuda_16": (i, [vp, i]),
        "fmvd_set_chunk_planes_cuda_8": (i, [vp, i]),
        "fmvd_deconvolve_volume_device_cuda_16": (i, [vp, C.POINTER(vp), i, i, vp, vp, vp, vp]),
        "fmvd_deconvolve_volume_device_cuda_8": (i, [vp, C.POINTER(vp), i, i, vp, vp, vp, vp]),
        "fmvd_deconvolve_volume_device_noalloc_cuda_16": (i, [vp, C.POINTER(vp), i, i, vp, vp, vp, vp]),
        "fmvd_deconvolve_volume_device_noalloc_cuda_8": (i, [vp, C.POINTER(vp), i, i, vp, vp, vp, vp]),
        "fmvd_set_autotune_cache": (None, [cp]),
        "fmvd_autotune_cache_selftest": (i, []),
        "fmvd_set_lookahead_cuda_16": (i, [vp, i]),
        "fmvd_set_lookahead_cuda_8": (i, [vp, i]),
        "fmvd_set_weights_device_cuda_16": (i, [vp, C.POINTER(vp), vp]),
        "fmvd_set_weights_device_cuda_8": (i, [vp, C.POINTER(vp), vp]),
        "fmvd_reconstruct_device_cuda_16": (i, [vp, C.POINTER(vp), i, i, i, C.POINTER(C.POINTER(f)), i, f,
                                                C.POINTER(vp), i, i, vp, vp, vp]),
        "fmvd_reconstruct_device_cuda_8": (i, [vp, C.POINTER(vp), i, i, i, C.POINTER(C.POINTER(f)), i, f,
                                               C.POINTER(vp), i, i, vp, vp, vp]),
        "fmvd_plan_info_cuda_16": (i, [vp, C.POINTER(C.c_int)]),
        "fmvd_plan_info_cuda_8": (i, [vp, C.POINTER(C.c_int)]),
        "fmvd_snap_transform_size": (i, [i]),
        "fmvd_fft_extent_supported": (i, [i, i]),
        "fmvd_kernel_timing_enable": (None, [i]),
        "fmvd_kernel_timing_reset": (None, []),
        "fmvd_kernel_timing_read": (i, [i, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
        "transform_cuda_16": (None, [vp, i, i, i, i, i, i, C.POINTER(f), cp, i, i, f, cp]),
        "transform_cuda_8": (None, [vp, i, i, i, i, i, i, C.POINTER(f), cp, i, i, f, cp]),
        "fmvd_transform_device_16": (i, [vp, i, i, i, vp, i, i, i, C.POINTER(f), vp]),
        "fmvd_transform_device_8": (i, [vp, i, i, i, vp, i, i, i, C.POINTER(f), vp]),
        "fmvd_transform_mask_device": (i, [i, i, i, i, i, i, C.POINTER(f), i, f, vp, vp]),
        "fmvd_autotune_cuda_16": (i, [vp, i, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
        "fmvd_autotune_cuda_8": (i, [vp, i, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
        "fmvd_set_variants_cuda_16": (i, [vp, C.POINTER(C.c_int)]),
        "fmvd_set_variants_cuda_8": (i, [vp, C.POINTER(C.c_int)]),
        "fmvd_convergence_cuda_16": (i, [vp, C.POINTER(vp), i, i, vp, vp, C.POINTER(C.c_double), vp]),
        "fmvd_convergence_cuda_8": (i, [vp, C.POINTER(vp), i, i, vp, vp, C.POINTER(C.c_double), vp]),
        "fmvd_device_malloc": (vp, [C.c_size_t]),
        "fmvd_device_free": (None, [vp]),
        "fmvd_memcpy": (i, [vp, vp, C.c_size_t, i]),
        "fmvd_memset_device": (i, [vp, i, C.c_size_t]),
        "fmvd_device_synchronize": (i, []),
        "fmvd_stream_create": (vp, []),
        "fmvd_stream_destroy": (i, [vp]),
        "fmvd_stream_synchronize": (i, [vp]),
        "fmvd_device_memory": (i, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    L.setErrorHandler(_error_cb, None)
    _lib = L
    return L


def exported_symbols() -> list:
    return ["setErrorHandler", "iDivUp", "getNumCUDADevices", "getCudaDeviceName", "setCudaDevice",
            "fmvd_malloc", "fmvd_free", "fmvd_initialize_cuda_16", "fmvd_initialize_cuda_8",
            "fmvd_deconvolve_planes_cuda_16", "fmvd_deconvolve_planes_cuda_8",
            "fmvd_destroy_cuda_16", "fmvd_destroy_cuda_8", "fmvd_deconvolve_files_cuda_16",
            "fmvd_deconvolve_files_cuda_8", "transform_cuda_16", "transform_cuda_8"]


def restore_error_handler() -> None:
    """Re-install the recording handler (JNI entry points install their own)."""
    lib().setErrorHandler(_error_cb, None)


def take_errors() -> list:
    out = list(_errors)
    _errors.clear()
    return out


def check():
    """Raise FmvdError if the library reported an error since the last call."""
    errs = take_errors()
    if errs:
        raise FmvdError("; ".join(errs))


def snap_transform_size(n: int) -> int:
    return lib().fmvd_snap_transform_size(int(n))


def fft_extent_supported(ffth: int, fftw: int) -> bool:
    return bool(lib().fmvd_fft_extent_supported(int(ffth), int(fftw)))


def _float_ptr_array(arrays: Sequence[np.ndarray]):
    arr_t = C.POINTER(C.c_float) * len(arrays)
    return arr_t(*[a.ctypes.data_as(C.POINTER(C.c_float)) for a in arrays])


class Deconvolver:
    """Mirror of the plan API (fmvd_initialize / fmvd_deconvolve_planes /
    fmvd_destroy) with in-memory plane streaming.

    kernels: V arrays (kh, kw) float32 -- normalised IN PLACE, as the C API does.
    weights: V arrays (dh, dw) float32.
    """

    def __init__(self, data_h: int, data_w: int, weights: Sequence[np.ndarray],
                 kernels: Sequence[np.ndarray], psf_type: int = 0, bits: int = 16, nstreams: int = 3,
                 batch_planes: int = 0, chunk_planes: int = 0):
        L = lib()
        self.bits = bits
        self.h, self.w = int(data_h), int(data_w)
        self.nviews = len(kernels)
        self.kernels = [np.ascontiguousarray(k, dtype=np.float32) for k in kernels]
        self.weights = None if weights is None else [np.ascontiguousarray(w, dtype=np.float32) for w in weights]
        # callback sequence of the last run: ("get", ok) and ("put", z) in call order
        self.calls: list = []
        kh, kw = self.kernels[0].shape
        self._dtype = np.uint16 if bits == 16 else np.uint8
        self._ctype = C.c_uint16 if bits == 16 else C.c_uint8
        self._planes_in: list = []
        self._planes_out: list = []
        self._est_out: list = []
        self._pos = 0
        plane = self.h * self.w

        def source(bufs, offset):
            if self._pos >= len(self._planes_in):
                self.calls.append(("get", 0))
                return 0
            self.calls.append(("get", 1))
            views = self._planes_in[self._pos]
            for v in range(self.nviews):
                dst = np.ctypeslib.as_array(C.cast(bufs[v], C.POINTER(self._ctype)), shape=(offset + plane,))
                dst[offset:offset + plane] = views[v].reshape(-1)
            self._pos += 1
            return 1

        def sink(buf):
            self.calls.append(("put", len(self._planes_out)))
            self._planes_out.append(np.ctypeslib.as_array(buf, shape=(plane,)).copy().reshape(self.h, self.w))

        def esink(buf, z, _user):
            self._est_out.append(np.ctypeslib.as_array(buf, shape=(plane,)).copy().reshape(self.h, self.w))

        if bits == 16:
            self._src, self._snk = SOURCE16_T(source), SINK16_T(sink)
            init, self._run, self._destroy = L.fmvd_initialize_cuda_16, L.fmvd_deconvolve_planes_cuda_16, L.fmvd_destroy_cuda_16
            self._set_esink, self._set_batch = L.fmvd_set_estimate_sink_cuda_16, L.fmvd_set_batch_planes_cuda_16
            self._set_chunk = L.fmvd_set_chunk_planes_cuda_16
            self._volume, self._info = L.fmvd_deconvolve_volume_device_cuda_16, L.fmvd_plan_info_cuda_16
        elif bits == 8:
            self._src, self._snk = SOURCE8_T(source), SINK8_T(sink)
            init, self._run, self._destroy = L.fmvd_initialize_cuda_8, L.fmvd_deconvolve_planes_cuda_8, L.fmvd_destroy_cuda_8
            self._set_esink, self._set_batch = L.fmvd_set_estimate_sink_cuda_8, L.fmvd_set_batch_planes_cuda_8
            self._set_chunk = L.fmvd_set_chunk_planes_cuda_8
            self._volume, self._info = L.fmvd_deconvolve_volume_device_cuda_8, L.fmvd_plan_info_cuda_8
        else:
            raise ValueError("bits must be 8 or 16")
        self._esink = ESINK_T(esink)
        take_errors()
        wptr = None if self.weights is None else _float_ptr_array(self.weights)
        self.plan = init(self.h, self.w, wptr, _float_ptr_array(self.kernels),
                         kh, kw, int(psf_type), self.nviews, int(nstreams), self._src, self._snk)
        check()
        if not self.plan:
            raise FmvdError("plan creation failed")
        if batch_planes:
            self._set_batch(self.plan, int(batch_planes))
        if chunk_planes:
            self._set_chunk(self.plan, int(chunk_planes))

    def info(self) -> dict:
        d = (C.c_int * 7)()
        self._info(self.plan, d)
        return dict(data_h=d[0], data_w=d[1], fft_h=d[2], fft_w=d[3], views=d[4], batch=d[5], chunk=d[6])

    def set_batch_planes(self, n: int) -> int:
        return self._set_batch(self.plan, int(n))

    def set_lookahead(self, n: int) -> int:
        """At most n planes pulled ahead of the sink, in the reference's call
        order (0 = batched pipeline); returns the value in effect."""
        L = lib()
        fn = L.fmvd_set_lookahead_cuda_16 if self.bits == 16 else L.fmvd_set_lookahead_cuda_8
        return fn(self.plan, int(n))

    def set_weights_device(self, weight_ptrs: Sequence[int], stream: int = 0) -> None:
        """Weights from device images (dataH x dataW float each), normalised on the device."""
        L = lib()
        fn = L.fmvd_set_weights_device_cuda_16 if self.bits == 16 else L.fmvd_set_weights_device_cuda_8
        arr = (C.c_void_p * self.nviews)(*[C.c_void_p(int(p)) for p in weight_ptrs])
        rc = fn(self.plan, arr, C.c_void_p(stream or None))
        check()
        if rc != 0:
            raise FmvdError("setting device weights failed")

    def reconstruct_device(self, acquired_ptrs: Sequence[int], src_shape, inverses, border: int, zspacing: float,
                           registered_ptrs: Sequence[int], nplanes: int, iterations: int, out_ptr: int = 0,
                           est_ptr: int = 0, stream: int = 0) -> None:
        """Transform -> border-mask weights -> RL on the device (src_shape = (d, h, w))."""
        L = lib()
        fn = L.fmvd_reconstruct_device_cuda_16 if self.bits == 16 else L.fmvd_reconstruct_device_cuda_8
        d, h, w = src_shape
        mats = [np.ascontiguousarray(m, dtype=np.float32).reshape(12) for m in inverses]
        inv = _float_ptr_array(mats)
        acq = (C.c_void_p * self.nviews)(*[C.c_void_p(int(p)) for p in acquired_ptrs])
        reg = (C.c_void_p * self.nviews)(*[C.c_void_p(int(p)) for p in registered_ptrs])
        rc = fn(self.plan, acq, int(w), int(h), int(d), inv, int(border), float(zspacing), reg, int(nplanes),
                int(iterations), C.c_void_p(out_ptr or None), C.c_void_p(est_ptr or None), C.c_void_p(stream or None))
        check()
        if rc != 0:
            raise FmvdError("device reconstruction failed")

    def set_chunk_planes(self, n: int) -> int:
        """Planes per cache-resident chunk (0 = automatic); returns the value in effect."""
        return self._set_chunk(self.plan, int(n))

    KINDS = ("column", "ratio", "update", "ingest", "egress")

    def autotune(self, planes: int = 32) -> dict:
        """Times every compiled FFT-plan variant per kernel kind and keeps the
        fastest.  Returns {kind: (chosen, [ms per variant or None])}."""
        L = lib()
        fn = L.fmvd_autotune_cuda_16 if self.bits == 16 else L.fmvd_autotune_cuda_8
        chosen = (C.c_int * 5)()
        times = (C.c_float * 15)()
        rc = fn(self.plan, int(planes), chosen, times)
        check()
        if rc != 0:
            raise FmvdError("autotune failed")
        return {k: (chosen[i], [times[3 * i + v] if times[3 * i + v] >= 0 else None for v in range(3)])
                for i, k in enumerate(self.KINDS)}

    def set_variants(self, variants) -> None:
        L = lib()
        fn = L.fmvd_set_variants_cuda_16 if self.bits == 16 else L.fmvd_set_variants_cuda_8
        arr = (C.c_int * 5)(*[int(x) for x in variants])
        if fn(self.plan, arr) != 0:
            raise FmvdError(f"invalid variants {list(variants)}")

    def run(self, planes: Sequence[Sequence[np.ndarray]], iterations: int, want_estimate: bool = False):
        """planes: sequence over z of per-view (dh, dw) sample arrays.
        Returns (samples[z, dh, dw], estimate[z, dh, dw] or None)."""
        self._planes_in = [[np.ascontiguousarray(v, dtype=self._dtype) for v in p] for p in planes]
        self._planes_out, self._est_out, self._pos, self.calls = [], [], 0, []
        self._set_esink(self.plan, self._esink if want_estimate else ESINK_T(), None)
        self._run(self.plan, int(iterations))
        check()
        out = np.stack(self._planes_out) if self._planes_out else np.zeros((0, self.h, self.w), self._dtype)
        est = np.stack(self._est_out) if want_estimate and self._est_out else None
        return out, est

    def run_device(self, view_ptrs: Sequence[int], nplanes: int, iterations: int, out_ptr: int = 0,
                   est_ptr: int = 0, init_ptr: int = 0, stream: int = 0) -> None:
        """GPU-resident volume: raw device pointers (ints), enqueued on `stream`."""
        arr = (C.c_void_p * self.nviews)(*[C.c_void_p(int(p)) for p in view_ptrs])
        rc = self._volume(self.plan, arr, int(nplanes), int(iterations), C.c_void_p(init_ptr or None),
                          C.c_void_p(out_ptr or None), C.c_void_p(est_ptr or None), C.c_void_p(stream or None))
        check()
        if rc != 0:
            raise FmvdError("volume deconvolution failed")

    def run_device_noalloc(self, view_ptrs: Sequence[int], nplanes: int, iterations: int, out_ptr: int = 0,
                           est_ptr: int = 0, init_ptr: int = 0, stream: int = 0) -> None:
        """run_device through the allocation-free entry point (work buffers in
        the library's fixed device scratch)."""
        L = lib()
        fn = (L.fmvd_deconvolve_volume_device_noalloc_cuda_16 if self.bits == 16
              else L.fmvd_deconvolve_volume_device_noalloc_cuda_8)
        arr = (C.c_void_p * self.nviews)(*[C.c_void_p(int(p)) for p in view_ptrs])
        rc = fn(self.plan, arr, int(nplanes), int(iterations), C.c_void_p(init_ptr or None),
                C.c_void_p(out_ptr or None), C.c_void_p(est_ptr or None), C.c_void_p(stream or None))
        check()
        if rc != 0:
            raise FmvdError("allocation-free volume deconvolution failed")

    def convergence(self, view_ptrs: Sequence[int], nplanes: int, iterations: int, ref_ptr: int,
                    init_ptr: int = 0, stream: int = 0) -> np.ndarray:
        """Relative L2 distance to the reference volume after ingest and after
        every iteration (FP64 sums on the device): iterations + 1 values."""
        L = lib()
        fn = L.fmvd_convergence_cuda_16 if self.bits == 16 else L.fmvd_convergence_cuda_8
        arr = (C.c_void_p * self.nviews)(*[C.c_void_p(int(p)) for p in view_ptrs])
        errs = (C.c_double * (iterations + 1))()
        rc = fn(self.plan, arr, int(nplanes), int(iterations), C.c_void_p(init_ptr or None), C.c_void_p(ref_ptr),
                errs, C.c_void_p(stream or None))
        check()
        if rc != 0:
            raise FmvdError("convergence run failed")
        return np.array(errs[:])

    def close(self):
        if getattr(self, "plan", None):
            self._destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def transform_to_file(volume: np.ndarray, inv: np.ndarray, target_shape, outfile: str,
                      mask_file: Optional[str] = None, border: int = 0, zspacing: float = 1.0) -> None:
    """Mirror of transform_cuda_{8,16}: volume[z, y, x] (uint8/uint16) -> raw file."""
    L = lib()
    vol = np.ascontiguousarray(volume)
    d, h, w = vol.shape
    td, th, tw = target_shape
    if vol.dtype == np.uint16:
        fn, ct = L.transform_cuda_16, C.c_uint16
    elif vol.dtype == np.uint8:
        fn, ct = L.transform_cuda_8, C.c_uint8
    else:
        raise TypeError("volume must be uint8 or uint16")
    planes = (C.POINTER(ct) * d)(*[vol[z].ctypes.data_as(C.POINTER(ct)) for z in range(d)])
    m = np.ascontiguousarray(inv, dtype=np.float32).reshape(12)
    take_errors()
    fn(planes, w, h, d, tw, th, td, m.ctypes.data_as(C.POINTER(C.c_float)), outfile.encode(),
       1 if mask_file else 0, int(border), float(zspacing), mask_file.encode() if mask_file else None)
    check()


def transform_device(src_ptr: int, src_shape, dst_ptr: int, target_shape, inv: np.ndarray, bits: int = 16,
                     stream: int = 0) -> None:
    L = lib()
    d, h, w = src_shape
    td, th, tw = target_shape
    m = np.ascontiguousarray(inv, dtype=np.float32).reshape(12)
    fn = L.fmvd_transform_device_16 if bits == 16 else L.fmvd_transform_device_8
    rc = fn(C.c_void_p(src_ptr), w, h, d, C.c_void_p(dst_ptr), tw, th, td,
            m.ctypes.data_as(C.POINTER(C.c_float)), C.c_void_p(stream or None))
    check()
    if rc != 0:
        raise FmvdError("transform failed")


def mask_device(src_shape, target_shape, inv: np.ndarray, border: int, zspacing: float, dst_ptr: int,
                stream: int = 0) -> None:
    L = lib()
    d, h, w = src_shape
    td, th, tw = target_shape
    m = np.ascontiguousarray(inv, dtype=np.float32).reshape(12)
    rc = L.fmvd_transform_mask_device(w, h, d, tw, th, td, m.ctypes.data_as(C.POINTER(C.c_float)), int(border),
                                      float(zspacing), C.c_void_p(dst_ptr), C.c_void_p(stream or None))
    check()
    if rc != 0:
        raise FmvdError("mask failed")


class KernelTiming:
    """Per-kernel HIP-event timing recorded by the library around launches."""

    def __enter__(self):
        lib().fmvd_kernel_timing_reset()
        lib().fmvd_kernel_timing_enable(1)
        return self

    def __exit__(self, *a):
        lib().fmvd_kernel_timing_enable(0)

    @staticmethod
    def read(kind: str):
        ms, n = C.c_double(), C.c_longlong()
        lib().fmvd_kernel_timing_read(TIMING_KINDS[kind], C.byref(ms), C.byref(n))
        return ms.value, n.value


class DeviceArray:
    """A device buffer owned by the library's HIP runtime (fmvd_device.h).

    Callers that also load another HIP runtime in the same process (a torch
    wheel bundles its own) cannot share device pointers with this library,
    so the tests and the benchmark allocate through here."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        take_errors()
        self.ptr = lib().fmvd_device_malloc(self.nbytes) or 0
        check()
        if not self.ptr:
            raise FmvdError("device allocation failed")

    @classmethod
    def from_host(cls, arr: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(arr)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        lib().fmvd_memcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes, 1)
        check()

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        lib().fmvd_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, 2)
        check()
        return out

    def download_planes(self, z0: int, nz: int) -> np.ndarray:
        """Planes [z0, z0 + nz) along the first axis."""
        per = int(np.prod(self.shape[1:], dtype=np.int64)) * self.dtype.itemsize
        assert 0 <= z0 and nz >= 0 and z0 + nz <= self.shape[0]
        out = np.empty((nz,) + self.shape[1:], self.dtype)
        lib().fmvd_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + z0 * per), nz * per, 2)
        check()
        return out

    def fill_zero(self) -> None:
        lib().fmvd_memset_device(C.c_void_p(self.ptr), 0, self.nbytes)
        check()

    def free(self) -> None:
        if self.ptr:
            lib().fmvd_device_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream of the library's runtime."""

    def __init__(self):
        self.handle = lib().fmvd_stream_create() or 0
        check()

    def synchronize(self) -> None:
        lib().fmvd_stream_synchronize(C.c_void_p(self.handle))
        check()

    def close(self) -> None:
        if self.handle:
            lib().fmvd_stream_destroy(C.c_void_p(self.handle))
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_synchronize() -> None:
    lib().fmvd_device_synchronize()
    check()


def device_memory():
    f, t = C.c_size_t(), C.c_size_t()
    lib().fmvd_device_memory(C.byref(f), C.byref(t))
    check()
    return f.value, t.value


class BatchVolume(C.Structure):
    """struct fmvd_batch_volume (fmvd_batch.h)."""
    _fields_ = [("dataH", C.c_int), ("dataW", C.c_int), ("nplanes", C.c_int), ("nViews", C.c_int),
                ("kernelH", C.c_int), ("kernelW", C.c_int), ("psf_type", C.c_int),
                ("h_weights", C.POINTER(C.POINTER(C.c_float))), ("h_kernels", C.POINTER(C.POINTER(C.c_float))),
                ("d_views", C.POINTER(C.c_void_p)), ("d_out", C.c_void_p), ("d_est", C.c_void_p)]


class Batch:
    """Mirror of the batched API (fmvd_batch.h): many independent volumes of
    uneven size, one graph launch per RL iteration.

    volumes: sequence of dicts with keys
      views   -- list of V DeviceArrays (Z, H, W) of samples
      weights -- list of V float32 arrays (H, W)
      kernels -- list of V float32 arrays (kh, kw), normalised IN PLACE
      psf_type (default 0), out (DeviceArray or None), est (DeviceArray or None)
    """

    def __init__(self, volumes, bits: int = 16):
        L = lib()
        if bits not in (8, 16):
            raise ValueError("bits must be 8 or 16")
        sfx = "16" if bits == 16 else "8"
        self._create = getattr(L, "fmvd_batch_create_cuda_" + sfx)
        self._run = getattr(L, "fmvd_batch_run_cuda_" + sfx)
        self._tune = getattr(L, "fmvd_batch_autotune_cuda_" + sfx)
        self._info = getattr(L, "fmvd_batch_info_cuda_" + sfx)
        self._destroy = getattr(L, "fmvd_batch_destroy_cuda_" + sfx)
        self._create.restype = C.c_void_p
        self._create.argtypes = [C.c_int, C.POINTER(BatchVolume)]
        self._run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        self._tune.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        self._info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        self._destroy.argtypes = [C.c_void_p]
        self._keep = []
        arr = (BatchVolume * len(volumes))()
        for i, v in enumerate(volumes):
            ks = [np.ascontiguousarray(k, dtype=np.float32) for k in v["kernels"]]
            ws = [np.ascontiguousarray(w, dtype=np.float32) for w in v["weights"]]
            nv = len(ks)
            z, h, w = v["views"][0].shape
            dv = (C.c_void_p * nv)(*[C.c_void_p(d.ptr) for d in v["views"]])
            kp, wp = _float_ptr_array(ks), _float_ptr_array(ws)
            self._keep += [ks, ws, dv, kp, wp]
            arr[i] = BatchVolume(h, w, z, nv, ks[0].shape[0], ks[0].shape[1], int(v.get("psf_type", 0)), wp, kp, dv,
                                 C.c_void_p(v["out"].ptr if v.get("out") is not None else None),
                                 C.c_void_p(v["est"].ptr if v.get("est") is not None else None))
            v["kernels_normalised"] = ks
        self._keep.append(arr)
        take_errors()
        self.handle = self._create(len(volumes), arr)
        check()
        if not self.handle:
            raise FmvdError("batch creation failed")

    def info(self):
        n = C.c_int()
        self._info(self.handle, C.byref(n), None, None)
        ext = (C.c_int * (2 * n.value))()
        pl = (C.c_int * n.value)()
        self._info(self.handle, C.byref(n), ext, pl)
        check()
        return [dict(fft_h=ext[2 * c], fft_w=ext[2 * c + 1], planes=pl[c]) for c in range(n.value)]

    def autotune(self):
        n = len(self.info())
        chosen = (C.c_int * (3 * n))()
        if self._tune(self.handle, chosen) != 0:
            check()
        check()
        return [tuple(chosen[3 * c:3 * c + 3]) for c in range(n)]

    def run(self, iterations: int, use_graph: bool = True, stream: int = 0) -> None:
        if self._run(self.handle, int(iterations), 1 if use_graph else 0, C.c_void_p(stream or None)) != 0:
            check()
        check()

    def close(self):
        if getattr(self, "handle", None):
            self._destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Net:
    """Mirror of the warm-start network API (fmvd_net.h)."""

    def __init__(self, seed: int = 0):
        L = lib()
        vp, i, f = C.c_void_p, C.c_int, C.c_float
        sigs = {
            "fmvd_net_param_count": (C.c_size_t, []),
            "fmvd_net_create": (vp, [C.c_ulonglong]),
            "fmvd_net_destroy": (None, [vp]),
            "fmvd_net_get_params": (i, [vp, vp]),
            "fmvd_net_set_params": (i, [vp, vp]),
            "fmvd_net_get_grads": (i, [vp, vp]),
            "fmvd_net_forward": (i, [vp, vp, i, i, i, i, vp, vp]),
            "fmvd_net_backward": (i, [vp, vp, vp, i, i, i, i, C.POINTER(C.c_double), vp]),
            "fmvd_net_adam": (i, [vp, f, vp]),
            "fmvd_net_prep_input": (i, [vp, i, i, i, i, i, vp, vp, vp]),
            "fmvd_net_infer": (i, [vp, vp, i, i, i, i, f, f, vp, vp]),
            "fmvd_net_infer_tiled": (i, [vp, vp, i, i, i, i, f, f, i, i, i, i, vp, vp]),
            "fmvd_net_mfma_layout_test": (i, [vp, vp, vp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        self.L = L
        self.nparams = int(L.fmvd_net_param_count())
        take_errors()
        self.handle = L.fmvd_net_create(int(seed))
        check()

    def _rc(self, rc):
        check()
        if rc != 0:
            raise FmvdError("network call failed")

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        self._rc(self.L.fmvd_net_get_params(self.handle, out.ctypes.data))
        return out

    def set_params(self, p: np.ndarray) -> None:
        a = np.ascontiguousarray(p, np.float32)
        assert a.size == self.nparams
        self._rc(self.L.fmvd_net_set_params(self.handle, a.ctypes.data))

    def grads(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        self._rc(self.L.fmvd_net_get_grads(self.handle, out.ctypes.data))
        return out

    def forward(self, d_x: "DeviceArray", shape, d_y: "DeviceArray", stream: int = 0) -> None:
        N, D, H, W = shape
        self._rc(self.L.fmvd_net_forward(self.handle, d_x.ptr, N, D, H, W, d_y.ptr, stream or None))

    def backward(self, d_x: "DeviceArray", d_t: "DeviceArray", shape, stream: int = 0) -> float:
        N, D, H, W = shape
        loss = C.c_double()
        self._rc(self.L.fmvd_net_backward(self.handle, d_x.ptr, d_t.ptr, N, D, H, W, C.byref(loss), stream or None))
        return loss.value

    def adam(self, lr: float, stream: int = 0) -> None:
        self._rc(self.L.fmvd_net_adam(self.handle, float(lr), stream or None))

    def infer(self, views, scale: float, d_out: "DeviceArray", floor_value: float = 1e-3, tile=None,
              stream: int = 0) -> None:
        """views: list of DeviceArrays (D, H, W) uint16; d_out float (D, H, W)."""
        D, H, W = views[0].shape
        ptrs = DeviceArray((len(views),), np.uint64)
        ptrs.upload(np.array([v.ptr for v in views], np.uint64))
        if tile is None:
            rc = self.L.fmvd_net_infer(self.handle, ptrs.ptr, len(views), D, H, W, float(scale), float(floor_value),
                                       d_out.ptr, stream or None)
        else:
            tz, ty, tx, halo = tile
            rc = self.L.fmvd_net_infer_tiled(self.handle, ptrs.ptr, len(views), D, H, W, float(scale),
                                             float(floor_value), tz, ty, tx, halo, d_out.ptr, stream or None)
        self._rc(rc)
        device_synchronize()
        ptrs.free()

    def close(self):
        if getattr(self, "handle", None):
            self.L.fmvd_net_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_bf16_bits(a: np.ndarray) -> np.ndarray:
    """float -> bfloat16 bit patterns (round to nearest even), uint16."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)
