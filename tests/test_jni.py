"""JNI layer (reference fastspim_NativeSPIMReconstructionCuda.cpp:13-538)
driven through a fake JNIEnv (tests/jni_fake), since the image has no JDK.

CPU tests cover the error paths, which raise java.lang.RuntimeException with
the reference's messages and leave no session behind.  GPU tests run the
interactive session with the getNextPlane/returnNextPlane upcalls, the file
front end and the transforms, and require results identical to the C API's
on the same inputs."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

FAKE_DIR = os.path.join(ROOT, "tests", "jni_fake")
FAKE = os.path.join(FAKE_DIR, "libfakejni.so")
PREFIX = "Java_fastspim_NativeSPIMReconstructionCuda_"

P, I, F, U8 = C.c_void_p, C.c_int, C.c_float, C.c_uint8
NATIVE_SIGS = {
    "getNumCudaDevices": (I, [P, P]),
    "getCudaDeviceName": (P, [P, P, I]),
    "setCudaDevice": (None, [P, P, I]),
    "transform16": (None, [P, P, P, I, I, I, P, I, I, I, P, U8, I, F, P]),
    "transform8": (None, [P, P, P, I, I, I, P, I, I, I, P, U8, I, F, P]),
    "deconvolve": (None, [P, P, P, P, I, I, I, P, P, I, I, I, I, I, I]),
    "deconvolve_1init": (None, [P, P, I, I, I, P, P, I, I, I, I, I]),
    "deconvolve_1interactive8": (None, [P, P, I]),
    "deconvolve_1interactive16": (None, [P, P, I]),
    "deconvolve_1quit8": (None, [P, P]),
    "deconvolve_1quit16": (None, [P, P]),
}
FAKE_SIGS = {
    "fj_env": (P, []), "fj_class": (P, []), "fj_string": (P, [C.c_char_p]),
    "fj_object_array": (P, [I]), "fj_object_array_set": (None, [P, I, P]),
    "fj_short_array": (P, [I, P]), "fj_byte_array": (P, [I, P]), "fj_float_array": (P, [I, P]),
    "fj_pending_exception": (C.c_char_p, []), "fj_clear_exception": (None, []),
    "fj_unexpected_calls": (I, []), "fj_local_refs_deleted": (I, []), "fj_global_refs": (I, []),
    "fj_set_methods_present": (None, [I]), "fj_set_planes": (None, [I, I, I, I, P]),
    "fj_planes_served": (I, []), "fj_returned_count": (I, []), "fj_returned_plane": (P, [I]),
    "fj_reset": (None, []), "fj_string_value": (C.c_char_p, [P]),
    "fj_call_log": (C.c_char_p, []), "fj_set_short_view": (None, [I, I, I]),
}


class Jni:
    def __init__(self, fm):
        if not os.path.exists(FAKE):
            subprocess.run(["make", "-s", "-C", FAKE_DIR], check=True)
        self.f = C.CDLL(FAKE)
        for n, (r, a) in FAKE_SIGS.items():
            fn = getattr(self.f, n)
            fn.restype, fn.argtypes = r, a
        self.fm = fm
        self.L = fm.lib()
        for n, (r, a) in NATIVE_SIGS.items():
            fn = getattr(self.L, PREFIX + n)
            fn.restype, fn.argtypes = r, a
        self.env = self.f.fj_env()
        self.cls = self.f.fj_class()
        self._keep = []

    def call(self, name, *args):
        return getattr(self.L, PREFIX + name)(self.env, self.cls, *args)

    def string(self, s):
        return self.f.fj_string(str(s).encode())

    def strings(self, items):
        arr = self.f.fj_object_array(len(items))
        for i, s in enumerate(items):
            self.f.fj_object_array_set(arr, i, self.string(s))
        return arr

    def floats(self, a):
        a = np.ascontiguousarray(a, np.float32)
        return self.f.fj_float_array(a.size, a.ctypes.data)

    def planes(self, vol):
        """Java short[][] / byte[][] of a (d, h, w) volume."""
        arr = self.f.fj_object_array(vol.shape[0])
        mk = self.f.fj_short_array if vol.dtype == np.uint16 else self.f.fj_byte_array
        for z in range(vol.shape[0]):
            p = np.ascontiguousarray(vol[z])
            self.f.fj_object_array_set(arr, z, mk(p.size, p.ctypes.data))
        return arr

    def exception(self, clear=True):
        e = self.f.fj_pending_exception()
        if clear:
            self.f.fj_clear_exception()
        return e.decode() if e is not None else None


@pytest.fixture
def jni(fm):
    j = Jni(fm)
    j.f.fj_reset()
    yield j
    assert j.f.fj_unexpected_calls() == 0, "the binding called a JNI slot the fake does not implement"
    j.f.fj_reset()
    fm.restore_error_handler()
    fm.take_errors()


def write_f32(path, a):
    np.ascontiguousarray(a, np.float32).tofile(path)
    return str(path)


# ------------------------------------------------------------- CPU: errors
def test_unsupported_bit_depth(jni, tmp_path):
    names = jni.strings([tmp_path / "x"])
    jni.call("deconvolve", names, jni.string(tmp_path / "o"), 8, 8, 1, names, names, 3, 3, 0, 1, 1, 12)
    assert jni.exception() == "Unsupported bit depth"
    jni.call("deconvolve_1init", 8, 8, 1, names, names, 3, 3, 0, 1, 32)
    assert jni.exception() == "Unsupported bit depth"


@pytest.mark.parametrize("depth", [8, 16])
def test_quit_and_interactive_without_session(jni, depth):
    jni.call(f"deconvolve_1quit{depth}")
    assert jni.exception() == "deconvolution already quit"
    jni.call(f"deconvolve_1interactive{depth}", 5)
    assert jni.exception() == "deconvolution not initialized"


def test_init_missing_and_short_files(jni, tmp_path):
    k = write_f32(tmp_path / "k0.raw", np.ones((3, 3)))
    w = write_f32(tmp_path / "w0.raw", np.ones((8, 8)))
    short = write_f32(tmp_path / "short.raw", np.ones(5))
    missing = str(tmp_path / "nope.raw")
    jni.call("deconvolve_1init", 8, 8, 1, jni.strings([w]), jni.strings([missing]), 3, 3, 0, 1, 16)
    assert jni.exception() == f"Could not open {missing} for reading"
    jni.call("deconvolve_1init", 8, 8, 1, jni.strings([short]), jni.strings([k]), 3, 3, 0, 1, 16)
    assert jni.exception() == f"Short read from {short}"
    # fewer names than views
    jni.call("deconvolve_1init", 8, 8, 1, jni.strings([w]), jni.strings([k]), 3, 3, 0, 2, 16)
    assert "fewer than nViews" in jni.exception()
    # no session was created by any of these
    jni.call("deconvolve_1quit16")
    assert jni.exception() == "deconvolution already quit"
    # the String local references taken per file were released
    assert jni.f.fj_local_refs_deleted() >= 3


def test_init_without_callback_methods(jni, tmp_path):
    k = write_f32(tmp_path / "k0.raw", np.ones((3, 3)))
    w = write_f32(tmp_path / "w0.raw", np.ones((8, 8)))
    jni.f.fj_set_methods_present(0)
    jni.call("deconvolve_1init", 8, 8, 1, jni.strings([w]), jni.strings([k]), 3, 3, 0, 1, 16)
    assert jni.exception() == "java/lang/NoSuchMethodError"
    jni.call("deconvolve_1quit16")
    assert jni.exception() == "deconvolution already quit"
    assert jni.f.fj_global_refs() == 0


def test_transform_argument_checks(jni, tmp_path):
    vol = np.zeros((2, 4, 4), np.uint16)
    out = jni.string(tmp_path / "t.raw")
    jni.call("transform16", jni.planes(vol), 4, 4, 2, jni.floats(np.eye(3, 4)[:, :].ravel()[:8]), 4, 4, 2,
             out, 0, 0, 1.0, None)
    assert jni.exception() == "invMatrix must hold 12 values"
    jni.call("transform16", jni.planes(vol), 4, 4, 3, jni.floats(np.eye(3, 4).ravel()), 4, 4, 2,
             out, 0, 0, 1.0, None)
    assert jni.exception() == "data holds fewer than d planes"
    jni.call("transform16", jni.planes(vol), 5, 4, 2, jni.floats(np.eye(3, 4).ravel()), 4, 4, 2,
             out, 0, 0, 1.0, None)
    assert jni.exception() == "a plane holds fewer than w*h samples"
    jni.call("transform8", jni.planes(vol.astype(np.uint8)), 0, 4, 2, jni.floats(np.eye(3, 4).ravel()),
             4, 4, 2, out, 0, 0, 1.0, None)
    assert jni.exception() == "invalid source dimensions"
    assert not os.path.exists(tmp_path / "t.raw")


def test_device_errors_become_runtime_exceptions(jni, fm):
    """Without a usable GPU the device calls raise (never exit the JVM)."""
    n = jni.call("getNumCudaDevices")
    if n > 0:
        pytest.skip("a GPU is present")
    assert n == 0
    assert jni.exception().startswith("HIP error")
    jni.call("setCudaDevice", 0)
    assert jni.exception().startswith("HIP error")


# ------------------------------------------------------------- GPU
def _write_case(tmp_path, data, kernels, weights, bits):
    Z, V, dh, dw = data.shape
    dt = np.uint16 if bits == 16 else np.uint8
    views, kf, wf = [], [], []
    for v in range(V):
        p = tmp_path / f"view{v}.raw"
        np.ascontiguousarray(data[:, v].astype(dt)).tofile(p)
        views.append(str(p))
        kf.append(write_f32(tmp_path / f"k{v}.raw", kernels[v]))
        wf.append(write_f32(tmp_path / f"w{v}.raw", weights[v]))
    return views, kf, wf


def _c_api(fm, data, kernels, weights, psf, iters, bits):
    Z, V, dh, dw = data.shape
    dt = np.uint16 if bits == 16 else np.uint8
    with fm.Deconvolver(dh, dw, list(weights), [k.copy() for k in kernels], psf_type=psf, bits=bits) as dec:
        out, _ = dec.run([[data[z, v].astype(dt) for v in range(V)] for z in range(Z)], iters)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("bits,psf", [(16, 0), (8, 2), (16, 3)])
def test_interactive_session_matches_c_api(jni, fm, tmp_path, bits, psf):
    from gpu_util import make_case
    Z, V, dh, dw, k, iters = 5, 3, 90, 110, 15, 6
    data, kernels, weights = make_case(Z, V, dh, dw, k, k, bits, seed=41 + bits)
    dt = np.uint16 if bits == 16 else np.uint8
    _, kf, wf = _write_case(tmp_path, data, kernels, weights, bits)
    assert jni.call("getNumCudaDevices") >= 1 and jni.exception() is None
    name = jni.f.fj_string_value(jni.call("getCudaDeviceName", 0))
    assert jni.exception() is None and name
    jni.call("setCudaDevice", 0)
    assert jni.exception() is None
    jni.call("deconvolve_1init", dw, dh, Z, jni.strings(wf), jni.strings(kf), k, k, psf, V, bits)
    assert jni.exception() is None
    assert jni.f.fj_global_refs() == 1
    jni.call("deconvolve_1init", dw, dh, Z, jni.strings(wf), jni.strings(kf), k, k, psf, V, bits)
    assert jni.exception() == "deconvolution already initialized"
    expect = _c_api(fm, data, kernels, weights, psf, iters, bits)
    # two runs on the same session: the plan is reusable, as in Fiji
    for rep in range(2):
        planes = np.ascontiguousarray(data.astype(dt))
        jni.f.fj_set_planes(bits, Z, V, dh * dw, planes.ctypes.data)
        before = jni.f.fj_returned_count()
        log0 = len(jni.f.fj_call_log())
        jni.call(f"deconvolve_1interactive{bits}", iters)
        assert jni.exception() is None
        # the reference's upcall order with nStreams = min(dataD, 3) = 3
        # (fmvd_deconvolve.impl.cpp:216-245,288-294): 3 gets, (return, get)
        # pairs until a get yields null, then the remaining returns
        assert jni.f.fj_call_log().decode()[log0:] == "GGGRGRGRgRR"
        assert jni.f.fj_planes_served() == Z
        assert jni.f.fj_returned_count() - before == Z
        got = np.stack([np.ctypeslib.as_array(C.cast(jni.f.fj_returned_plane(before + z), C.POINTER(
            C.c_uint16 if bits == 16 else C.c_uint8)), shape=(dh * dw,)).reshape(dh, dw) for z in range(Z)])
        np.testing.assert_array_equal(got, expect)
    jni.call(f"deconvolve_1quit{bits}")
    assert jni.exception() is None
    assert jni.f.fj_global_refs() == 0
    jni.call(f"deconvolve_1quit{bits}")
    assert jni.exception() == "deconvolution already quit"


@pytest.mark.gpu
@pytest.mark.parametrize("bits", [16, 8])
def test_file_front_end_matches_c_api(jni, fm, tmp_path, bits):
    """JNI deconvolve() (fastspim_NativeSPIMReconstructionCuda.cpp:171-255)
    and the CLI (fmvd_deconvolve_hdd.cpp) write the C API's samples."""
    from gpu_util import make_case
    Z, V, dh, dw, k, iters, psf = 7, 2, 70, 64, 11, 5, 1
    data, kernels, weights = make_case(Z, V, dh, dw, k, k, bits, seed=5 + bits)
    dt = np.uint16 if bits == 16 else np.uint8
    views, kf, wf = _write_case(tmp_path, data, kernels, weights, bits)
    expect = _c_api(fm, data, kernels, weights, psf, iters, bits)
    out = tmp_path / "jni_out.raw"
    jni.call("deconvolve", jni.strings(views), jni.string(out), dw, dh, Z, jni.strings(wf), jni.strings(kf),
             k, k, psf, V, iters, bits)
    assert jni.exception() is None
    np.testing.assert_array_equal(np.fromfile(out, dt).reshape(Z, dh, dw), expect)

    cli = os.path.join(ROOT, "gpu_deconvolution_amd", "fmvd_deconvolve")
    cli_out = tmp_path / "cli_out.raw"
    args = [cli, "-v", str(V), "-i", str(iters), "-t", str(psf), "-o", str(cli_out), "-d", *views, "-w", *wf,
            "-k", *kf, "-dw", str(dw), "-dh", str(dh), "-dd", str(Z), "-kw", str(k), "-kh", str(k)]
    if bits == 8:
        args.append("-8bit")
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    np.testing.assert_array_equal(np.fromfile(cli_out, dt).reshape(Z, dh, dw), expect)


@pytest.mark.gpu
@pytest.mark.parametrize("bits", [16, 8])
def test_transform_matches_c_api(jni, fm, tmp_path, bits):
    rng = np.random.default_rng(bits)
    d, h, w = 9, 33, 47
    dt = np.uint16 if bits == 16 else np.uint8
    vol = rng.integers(0, (1 << bits) - 1, size=(d, h, w)).astype(dt)
    c, s = np.cos(0.3), np.sin(0.3)
    inv = np.array([[c, -s, 0.1, 2.5], [s, c, 0.0, -1.25], [0.05, 0.0, 0.9, 0.5]], np.float32)
    tw, th, td = 41, 37, 11
    out_c, mask_c = tmp_path / "c.raw", tmp_path / "c_mask.raw"
    fm.transform_to_file(vol, inv, (td, th, tw), str(out_c), mask_file=str(mask_c), border=6,
                         zspacing=2.0)
    out_j, mask_j = tmp_path / "j.raw", tmp_path / "j_mask.raw"
    jni.call(f"transform{bits}", jni.planes(vol), w, h, d, jni.floats(inv.ravel()), tw, th, td,
             jni.string(out_j), 1, 6, 2.0, jni.string(mask_j))
    assert jni.exception() is None
    assert open(out_j, "rb").read() == open(out_c, "rb").read()
    assert open(mask_j, "rb").read() == open(mask_c, "rb").read()


@pytest.mark.gpu
def test_interactive_short_view_raises(jni, fm, tmp_path):
    """A getNextPlane array whose view is shorter than dataW*dataH ends the
    stream with a RuntimeException; no upcall follows the exception."""
    from gpu_util import make_case
    Z, V, dh, dw, k = 5, 2, 40, 50, 9
    data, kernels, weights = make_case(Z, V, dh, dw, k, k, 16, seed=3)
    _, kf, wf = _write_case(tmp_path, data, kernels, weights, 16)
    jni.call("deconvolve_1init", dw, dh, Z, jni.strings(wf), jni.strings(kf), k, k, 0, V, 16)
    assert jni.exception() is None
    planes = np.ascontiguousarray(data)
    jni.f.fj_set_planes(16, Z, V, dh * dw, planes.ctypes.data)
    jni.f.fj_set_short_view(3, 1, dh * dw - 5)
    jni.call("deconvolve_1interactive16", 2)
    assert jni.exception() == "getNextPlane returned a view with fewer than planesize samples"
    assert jni.f.fj_call_log().decode() == "GGGRG"
    jni.call("deconvolve_1quit16")
    assert jni.exception() is None
