"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/*.h, and validates arguments before touching the device
(errors reach the installed handler; the default handler exits like the
reference's, fmvd_cuda_utils.cpp:12-16)."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:JNIEXPORT\s+)?[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", src, flags=re.M):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof", "JNICALL", "typedef", "int", "void",
                        "float", "char", "unsigned"):
                continue
            names.add(name)
    # typedef'd function pointer names are not symbols
    return {n for n in names if not n.endswith("_t") and not n.startswith("datasource") and not n.startswith("datasink")}


def test_every_declared_function_is_exported(fm):
    out = subprocess.run(["nm", "-D", "--defined-only", fm.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    decl = declared_functions()
    assert len(decl) >= 40
    missing = sorted(decl - exported)
    assert not missing, missing


def test_reference_entry_points_present(fm):
    L = fm.lib()
    for name in fm.exported_symbols():
        assert hasattr(L, name)
    jni = ["getNumCudaDevices", "getCudaDeviceName", "setCudaDevice", "transform8", "transform16", "deconvolve",
           "deconvolve_1quit8", "deconvolve_1quit16", "deconvolve_1interactive8", "deconvolve_1interactive16",
           "deconvolve_1init"]
    for j in jni:
        assert hasattr(L, "Java_fastspim_NativeSPIMReconstructionCuda_" + j)


def test_idivup_and_snap_match_oracle(fm, orc):
    L = fm.lib()
    for a, b in [(10, 3), (9, 3), (1, 256), (256, 256), (257, 256)]:
        assert L.iDivUp(a, b) == -(-a // b)
    for n in range(1, 5000, 7):
        assert fm.snap_transform_size(n) == orc.snap_transform_size(n)


def test_supported_extents_cover_workloads(fm):
    for h, w in [(512, 512), (1536, 1536), (2560, 2560), (256, 256), (1024, 1024), (256, 1024), (512, 256)]:
        assert fm.fft_extent_supported(h, w)
    # every extent snapTransformSize produces from 32 to 8192, on either axis
    sizes = sorted({fm.snap_transform_size(n) for n in range(1, 8193)})
    assert sizes[0] == 16 and sizes[-1] == 8192 and 4608 in sizes and 7680 in sizes
    for n in sizes:
        assert fm.fft_extent_supported(n, 512) and fm.fft_extent_supported(512, n), n
    assert not fm.fft_extent_supported(8704, 512)
    assert not fm.fft_extent_supported(512, 8704)


def _kernels(v=2, k=5):
    return [np.ones((k, k), np.float32) for _ in range(v)]


@pytest.mark.parametrize("kwargs,msg", [
    (dict(data_h=0, data_w=10), "invalid plan dimensions"),
    (dict(data_h=10, data_w=10, psf_type=7), "invalid PSF type"),
    (dict(data_h=10, data_w=9000), "unsupported FFT extent"),
])
def test_plan_validation_reports_through_handler(fm, kwargs, msg):
    h, w = max(kwargs["data_h"], 1), max(kwargs["data_w"], 1)
    with pytest.raises(fm.FmvdError, match=msg):
        fm.Deconvolver(kwargs["data_h"], kwargs["data_w"], [np.ones((h, w), np.float32)] * 2, _kernels(),
                       psf_type=kwargs.get("psf_type", 0))


def test_transform_validation(fm):
    with pytest.raises(fm.FmvdError, match="invalid transform dimensions"):
        fm.transform_device(0, (0, 4, 4), 0, (1, 4, 4), np.eye(3, 4))


def test_default_handler_exits_like_reference():
    code = ("import ctypes, numpy as np, gpu_deconvolution_amd as fm\n"
            "L = ctypes.CDLL(fm.LIB_PATH)\n"
            "L.fmvd_transform_device_16.argtypes = [ctypes.c_void_p] + [ctypes.c_int]*3 + [ctypes.c_void_p] + "
            "[ctypes.c_int]*3 + [ctypes.c_void_p, ctypes.c_void_p]\n"
            "m = np.zeros(12, np.float32)\n"
            "L.fmvd_transform_device_16(None, 0, 1, 1, None, 1, 1, 1, m.ctypes.data, None)\n"
            "print('not reached')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 1
    assert "exiting..." in r.stdout
    assert "not reached" not in r.stdout


def test_last_error_message(fm):
    import ctypes
    with pytest.raises(fm.FmvdError):
        fm.transform_device(0, (1, 1, 0), 0, (1, 1, 1), np.eye(3, 4))
    buf = ctypes.create_string_buffer(256)
    n = fm.lib().fmvd_last_error(buf, 256)
    assert n > 0 and b"invalid transform dimensions" in buf.value


def test_batch_rejects_bad_volumes_before_touching_the_device(fm):
    """fmvd_batch_create validates every volume first (no GPU needed)."""
    import ctypes as C
    L = fm.lib()
    L.fmvd_batch_create_cuda_16.restype = C.c_void_p
    L.fmvd_batch_create_cuda_16.argtypes = [C.c_int, C.POINTER(fm.BatchVolume)]
    w = np.ones((8, 8), np.float32)
    k = np.ones((3, 3), np.float32)
    wp, kp = fm._float_ptr_array([w]), fm._float_ptr_array([k])
    dv = (C.c_void_p * 1)(C.c_void_p(0x1000))
    vol = fm.BatchVolume(8, 8, 2, 1, 3, 3, 0, wp, kp, dv, None, None)   # FFT 16x16: unsupported
    arr = (fm.BatchVolume * 1)(vol)
    fm.take_errors()
    assert not L.fmvd_batch_create_cuda_16(1, arr)
    errs = fm.take_errors()
    assert errs and "unsupported FFT extent" in errs[0]
    arr[0].nViews = 0
    assert not L.fmvd_batch_create_cuda_16(1, arr)
    errs = fm.take_errors()
    assert errs and "invalid volume dimensions" in errs[0]


def test_autotune_cache_line_format(fm):
    assert fm.lib().fmvd_autotune_cache_selftest() == 0
