"""The allocation-free entry point fmvd_deconvolve_volume_device_noalloc_cuda_*
against fmvd_deconvolve_volume_device_cuda_* (bit for bit), twice in a row
on one plan, the second call with a smaller volume; at an FFT extent whose
planes do not all fit the fixed scratch at once (batches of 6 at 1536^2)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def volume(rng, Z, V, H, W, bits):
    hi = 4000 if bits == 16 else 250
    return [rng.integers(20, hi, size=(Z, H, W)).astype(np.uint16 if bits == 16 else np.uint8) for _ in range(V)]


def run_both(fm, dec, views, Z, iters, init=None):
    H, W = views[0].shape[1:]
    dt = views[0].dtype
    dv = [fm.DeviceArray.from_host(v) for v in views]
    di = fm.DeviceArray.from_host(init) if init is not None else None
    res = []
    for fn in (dec.run_device, dec.run_device_noalloc):
        out = fm.DeviceArray((Z, H, W), dt)
        est = fm.DeviceArray((Z, H, W), np.float32)
        fn([d.ptr for d in dv], Z, iters, out_ptr=out.ptr, est_ptr=est.ptr, init_ptr=di.ptr if di else 0)
        fm.device_synchronize()
        res.append((out.download(), est.download()))
        out.free()
        est.free()
    for d in dv + ([di] if di else []):
        d.free()
    return res


@pytest.mark.parametrize("bits", [16, 8])
def test_noalloc_matches_allocating_entry_point(fm, bits):
    rng = np.random.default_rng(bits)
    V, H, W, k = 2, 1100, 1000, 41   # FFT 1536 x 1536: the scratch holds 6 planes
    print(f"\n[noalloc] bits={bits} {H}x{W} views={V}", flush=True)
    kernels = [np.exp(-0.5 * ((np.arange(k)[:, None] - 20) ** 2 / (1.5 + v) ** 2 +
                              (np.arange(k)[None, :] - 20) ** 2 / 2.0 ** 2)).astype(np.float32) for v in range(V)]
    weights = [rng.uniform(0.2, 1.0, size=(H, W)).astype(np.float32) for _ in range(V)]
    with fm.Deconvolver(H, W, weights, [x.copy() for x in kernels], bits=bits) as dec:
        assert (dec.info()["fft_h"], dec.info()["fft_w"]) == (1536, 1536)
        # first call: 13 planes (three scratch batches), cold start
        (oa, ea), (ob, eb) = run_both(fm, dec, volume(rng, 13, V, H, W, bits), 13, 3)
        assert np.array_equal(oa, ob) and np.array_equal(ea, eb)
        # second call on the same plan: a smaller volume with a warm start
        init = rng.uniform(50, 500, size=(4, H, W)).astype(np.float32)
        (oa, ea), (ob, eb) = run_both(fm, dec, volume(rng, 4, V, H, W, bits), 4, 2, init=init)
        assert np.array_equal(oa, ob) and np.array_equal(ea, eb)


def test_noalloc_small_planes(fm):
    rng = np.random.default_rng(5)
    V, H, W, k = 3, 90, 70, 13
    kernels = [np.ones((k, k), np.float32) + np.eye(k, dtype=np.float32) * v for v in range(V)]
    weights = [rng.uniform(0.2, 1.0, size=(H, W)).astype(np.float32) for _ in range(V)]
    with fm.Deconvolver(H, W, weights, [x.copy() for x in kernels]) as dec:
        for Z in (40, 7):
            (oa, ea), (ob, eb) = run_both(fm, dec, volume(rng, Z, V, H, W, 16), Z, 4)
            assert np.array_equal(oa, ob) and np.array_equal(ea, eb)


def test_noalloc_rejects_planes_beyond_its_scratch(fm):
    """A plane whose spectrum and estimate exceed the fixed scratch (FFT
    8192 x 2560: 168 MB) is rejected with an error, before any launch."""
    H, W = 8150, 2500
    kernels = [np.ones((41, 41), np.float32) for _ in range(2)]
    with fm.Deconvolver(H, W, None, kernels) as dec:
        assert (dec.info()["fft_h"], dec.info()["fft_w"]) == (8192, 2560)
        d = fm.DeviceArray((1, H, W), np.uint16)
        with pytest.raises(fm.FmvdError, match="allocation-free"):
            dec.run_device_noalloc([d.ptr, d.ptr], 1, 1)
        d.free()
