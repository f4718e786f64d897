"""GPU-resident chain transform -> border-mask weights -> RL
(fmvd_reconstruct_device_cuda_*, fmvd_set_weights_device_cuda_*) against the
oracle's transform (fmvd_transform.impl.cu:24-56), mask
(fmvd_transform.cu:19-67) and RL (fmvd_deconvolve.impl.cpp:86-124,205-300),
which the reference chains through raw files."""
import numpy as np
import pytest

from gpu_util import check_rl, make_case

pytestmark = pytest.mark.gpu


def acquired_views(fm, shape, V, seed):
    from gpu_deconvolution_amd import synth
    vols = synth.gpu_views(shape, V, beads=60, seed=seed)
    host = [v.download() for v in vols]
    return vols, host


@pytest.mark.parametrize("bits", [16])
def test_reconstruct_device_matches_oracle(fm, orc, bits):
    from gpu_deconvolution_amd import synth
    V, src_shape, tshape, k, iters, border, zsp = 3, (22, 96, 104), (20, 90, 100), 15, 5, 8, 1.5
    print(f"\n[chain] V={V} source {src_shape} target {tshape}", flush=True)
    rng = np.random.default_rng(31)
    mats = [synth.random_affine(rng, src_shape, max_rot_deg=10.0, max_shift=3.0) for _ in range(V)]
    vols, host = acquired_views(fm, src_shape, V, seed=77)
    kernels = np.stack([synth.gaussian_kernel_2d(1.2 + 0.3 * v, 1.6, k) for v in range(V)])
    Z, H, W = tshape
    reg = [fm.DeviceArray(tshape, np.uint16) for _ in range(V)]
    out = fm.DeviceArray(tshape, np.uint16)
    est = fm.DeviceArray(tshape, np.float32)
    stream = fm.Stream()
    with fm.Deconvolver(H, W, None, [x.copy() for x in kernels], psf_type=0, bits=bits) as dec:
        dec.reconstruct_device([v.ptr for v in vols], src_shape, mats, border, zsp, [r.ptr for r in reg], Z, iters,
                               out_ptr=out.ptr, est_ptr=est.ptr, stream=stream.handle)
        stream.synchronize()
    reg_o = [orc.transform(host[v], mats[v], tshape) for v in range(V)]
    for v in range(V):
        assert np.array_equal(reg[v].download(), reg_o[v]), f"view {v} transform differs"
    w_o = np.stack([orc.mask(src_shape, tshape, mats[v], border, zsp) for v in range(V)])
    assert (w_o > 0).any()
    data = np.ascontiguousarray(np.stack(reg_o, axis=1))
    out_o, est_o = orc.rl_planes(data, w_o, kernels.copy(), 0, iters, threads=8)
    r, flips = check_rl(out.download(), est.download(), out_o, est_o)
    print(f"  relL2={r:.2e} flips={flips}", flush=True)
    for d in vols + reg + [out, est]:
        d.free()
    stream.close()


def test_device_weights_equal_host_weights(fm):
    """Weights normalised on the device are the host's, bit for bit: the same
    images given at init or through fmvd_set_weights_device give identical RL."""
    Z, V, dh, dw, k = 3, 3, 70, 66, 13
    data, kernels, weights = make_case(Z, V, dh, dw, k, k, 16, seed=8)
    weights[1, :10, :] = 0.0
    weights[:, 20:25, 30:40] = 0.0  # pixels where every view weighs 0
    planes = [[data[z, v] for v in range(V)] for z in range(Z)]
    with fm.Deconvolver(dh, dw, list(weights), [x.copy() for x in kernels]) as dec:
        out_h, est_h = dec.run(planes, 6, want_estimate=True)
    dws = [fm.DeviceArray.from_host(w) for w in weights]
    with fm.Deconvolver(dh, dw, None, [x.copy() for x in kernels]) as dec:
        dec.set_weights_device([d.ptr for d in dws])
        fm.device_synchronize()
        out_d, est_d = dec.run(planes, 6, want_estimate=True)
    assert np.array_equal(out_h, out_d)
    assert np.array_equal(est_h, est_d)
    for d in dws:
        d.free()
