"""GPU parity of the plane-wise multi-view RL (through the C ABI) against the
FP64 oracle: FP32 estimate within 1e-4 relative L2 of the FP64 estimate over
the data region, truncated samples equal except one-LSB flips where the FP64
estimate sits within 1e-2 of an integer.  Every FFT extent family the device
kernels instantiate for the workloads is covered (256 ... 3584, radix 2/3/5/7
and the 1280/1536/1792 row lengths), as are all PSF types, both bit depths,
non-square kernels and data, ragged batches, an empty source, the
GPU-resident volume API and the warm-start ingest."""
import numpy as np
import pytest

from gpu_util import check_rl, empty_device, from_device, make_case, to_device

pytestmark = pytest.mark.gpu


def run_gpu(fm, data, kernels, weights, psf, iters, bits, batch=0, nstreams=3, chunk=0):
    Z, V, dh, dw = data.shape
    dt = np.uint16 if bits == 16 else np.uint8
    with fm.Deconvolver(dh, dw, list(weights), [k.copy() for k in kernels], psf_type=psf, bits=bits,
                        nstreams=nstreams, batch_planes=batch, chunk_planes=chunk) as dec:
        planes = [[data[z, v].astype(dt) for v in range(V)] for z in range(Z)]
        out, est = dec.run(planes, iters, want_estimate=True)
        info = dec.info()
    return out, est, info


CASES = [
    # Z, V, dh, dw, kh, kw, psf, bits, iters   (fft extents in the comment)
    (4, 2, 64, 64, 15, 15, 0, 16, 10),        # 128 x 128
    (3, 3, 100, 120, 9, 13, 1, 16, 5),        # 128 x 256, non-square kernel
    (2, 2, 200, 180, 21, 21, 2, 8, 4),        # 256 x 256, 8-bit
    (2, 2, 230, 250, 31, 31, 3, 16, 4),       # 512 x 512
    (2, 2, 256, 256, 31, 31, 0, 16, 20),      # 512 x 512 (small workload's plane)
    (1, 2, 600, 500, 41, 41, 0, 16, 3),       # 1024 x 1024
    (1, 4, 1024, 1024, 41, 41, 0, 16, 2),     # 1536 x 1536 (real-size workload's plane)
    (1, 2, 2048, 700, 41, 41, 0, 16, 2),      # 2560 x 1024
    (1, 2, 300, 2048, 41, 41, 2, 16, 2),      # 512 x 2560
    (1, 2, 40, 3000, 41, 41, 0, 16, 2),       # 128 x 3072
    (1, 2, 3500, 60, 41, 41, 0, 16, 2),       # 3584 x 128
    (1, 2, 90, 1700, 31, 31, 0, 16, 2),       # 128 x 2048 (row length 1024)
    (1, 2, 90, 1750, 41, 41, 0, 8, 2),        # 128 x 2048
    (1, 2, 1990, 90, 41, 41, 0, 16, 2),       # 2048 x 128
    (1, 2, 70, 2600, 41, 41, 0, 16, 2),       # 128 x 3072
    (1, 2, 50, 3200, 41, 41, 0, 16, 2),       # 128 x 3584 (row length 1792)
    (1, 2, 50, 1200, 41, 41, 0, 16, 2),       # 128 x 1536 (row length 768)
    (1, 2, 50, 2300, 41, 41, 0, 16, 2),       # 128 x 2560 (row length 1280)
    (2, 2, 100, 90, 16, 12, 0, 16, 5),        # 128 x 128, even kernel: explicit kernel-hat spectrum
    (2, 2, 90, 100, 15, 12, 0, 8, 4),         # 128 x 128, odd x even kernel: explicit, 8-bit
    (2, 3, 101, 77, 15, 15, 0, 16, 5),        # 128 x 128, odd data width (per-sample pad map loads)
    (2, 2, 77, 131, 13, 17, 1, 8, 4),         # 128 x 256, odd width, 8-bit
    (2, 2, 7, 6, 9, 9, 0, 16, 5),             # 16 x 16: the smallest extent (row length 8)
    (2, 2, 300, 6, 9, 9, 0, 16, 5),           # 512 x 16
    (2, 2, 6, 300, 9, 9, 2, 16, 5),           # 16 x 512
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[:6])) + f"-psf{c[6]}-b{c[7]}")
def test_rl_parity(fm, orc, case):
    Z, V, dh, dw, kh, kw, psf, bits, iters = case
    print(f"\n[rl parity] case={case}", flush=True)
    data, kernels, weights = make_case(Z, V, dh, dw, kh, kw, bits, seed=dh * 7 + dw)
    out, est, info = run_gpu(fm, data, kernels, weights, psf, iters, bits)
    assert out.shape == (Z, dh, dw)
    out_o, est_o = orc.rl_planes(data, weights, kernels.copy(), psf, iters, bits=bits, threads=8)
    r, flips = check_rl(out, est, out_o, est_o, bits=bits)
    print(f"  fft={info['fft_h']}x{info['fft_w']} relL2={r:.2e} boundary flips={flips}", flush=True)


# Extents above 4096 (snapTransformSize's multiples of 512, fmvd_deconvolve.cpp:28-51):
# every row length N = fftW/2 of 2304 ... 4096 and every column length 4608 ...
# 8192, on thin planes so the oracle stays fast.
LARGE_CASES = [(50, w) for w in (4560, 5000, 5500, 6000, 6500, 7000, 7500, 8000)] + \
              [(h, 50) for h in (4560, 5000, 5500, 6000, 6500, 7000, 7500, 8000)]


@pytest.mark.parametrize("case", LARGE_CASES, ids=lambda c: f"{c[0]}x{c[1]}")
def test_rl_parity_large_extents(fm, orc, case):
    dh, dw = case
    print(f"\n[rl parity large] {dh}x{dw}", flush=True)
    data, kernels, weights = make_case(1, 2, dh, dw, 41, 41, 16, seed=dh + 3 * dw, beads=30)
    out, est, info = run_gpu(fm, data, kernels, weights, 0, 2, 16)
    assert max(info["fft_h"], info["fft_w"]) > 4096
    out_o, est_o = orc.rl_planes(data, weights, kernels.copy(), 0, 2, threads=8)
    r, flips = check_rl(out, est, out_o, est_o)
    print(f"  fft={info['fft_h']}x{info['fft_w']} relL2={r:.2e} boundary flips={flips}", flush=True)


VARIANT_CASES = [
    (1, 2, 230, 250, 31, 31),     # 512 x 512
    (1, 2, 600, 500, 41, 41),     # 1024 x 1024
    (1, 2, 1024, 1024, 41, 41),   # 1536 x 1536
    (1, 2, 2048, 300, 41, 41),    # 2560 x 512
    (1, 2, 90, 2300, 41, 41),     # 128 x 2560 (row length 1280)
]


@pytest.mark.parametrize("case", VARIANT_CASES, ids=lambda c: f"{c[2]}x{c[3]}")
def test_every_fft_plan_variant_matches_oracle(fm, orc, case):
    """Each compiled FFT-plan variant (the autotuner's search space) computes
    the same RL to the oracle's tolerance; the autotuner's pick is one of them."""
    Z, V, dh, dw, kh, kw = case
    data, kernels, weights = make_case(Z, V, dh, dw, kh, kw, 16, seed=dh + 3 * dw)
    out_o, est_o = orc.rl_planes(data, weights, kernels.copy(), 0, 2, threads=8)
    planes = [[data[z, v] for v in range(V)] for z in range(Z)]
    with fm.Deconvolver(dh, dw, list(weights), [k.copy() for k in kernels]) as dec:
        tuned = dec.autotune(planes=2)
        print(f"\n[variants] {case} tuned={ {k: c for k, (c, _) in tuned.items()} }", flush=True)
        for kind_idx, kind in enumerate(dec.KINDS):
            nvar = sum(t is not None for t in tuned[kind][1])
            for var in range(nvar):
                vs = [0] * 5
                vs[kind_idx] = var
                dec.set_variants(vs)
                if kind == "ingest":
                    continue  # warm-start ingest is covered by test_warm_start_matches_oracle
                out, est = dec.run(planes, 2, want_estimate=True)
                check_rl(out, est, out_o, est_o)
        dec.set_variants([c for c, _ in tuned.values()])
        out, est = dec.run(planes, 2, want_estimate=True)
        check_rl(out, est, out_o, est_o)


@pytest.mark.parametrize("dh, dw, k", [(64, 80, 15), (1024, 1024, 41)])
def test_conjugate_kernel_hat_matches_explicit(fm, orc, monkeypatch, dh, dw, k):
    """Independent PSFs of odd size: pass C multiplies by the conjugate of the
    kernel spectrum instead of a separate kernel-hat spectrum (the flipped
    kernel's).  FMVD_EXPLICIT_KHAT=1 builds the explicit one; both agree with
    each other to FP32 rounding and with the oracle."""
    Z, V = 2, 3
    data, kernels, weights = make_case(Z, V, dh, dw, k, k, 16, seed=dh + k)
    monkeypatch.delenv("FMVD_EXPLICIT_KHAT", raising=False)
    out_c, est_c, _ = run_gpu(fm, data, kernels, weights, 0, 4, 16)
    monkeypatch.setenv("FMVD_EXPLICIT_KHAT", "1")
    out_e, est_e, _ = run_gpu(fm, data, kernels, weights, 0, 4, 16)
    rel = float(np.linalg.norm(est_c - est_e) / np.linalg.norm(est_e))
    print(f"\n[khat] {dh}x{dw} k={k}: conj vs explicit relL2={rel:.2e}", flush=True)
    assert rel < 1e-5
    out_o, est_o = orc.rl_planes(data, weights, kernels.copy(), 0, 4, threads=8)
    check_rl(out_c, est_c, out_o, est_o)
    check_rl(out_e, est_e, out_o, est_o)


def test_batches_and_device_volume_agree(fm):
    Z, V, dh, dw, k = 7, 2, 90, 110, 11
    data, kernels, weights = make_case(Z, V, dh, dw, k, k, 16, seed=3)
    ref_out, ref_est, _ = run_gpu(fm, data, kernels, weights, 0, 6, 16, batch=1)
    for b, c in ((2, 0), (3, 0), (8, 0), (8, 3), (7, 2), (8, 1)):
        # c: planes per cache-resident chunk inside a batch (0 = automatic)
        out, est, _ = run_gpu(fm, data, kernels, weights, 0, 6, 16, batch=b, chunk=c)
        assert np.array_equal(out, ref_out) and np.array_equal(est, ref_est), (b, c)
    # GPU-resident volume through the same plan type
    dviews = [to_device(np.ascontiguousarray(data[:, v])) for v in range(V)]
    dout = empty_device((Z, dh, dw), np.uint16)
    dest = empty_device((Z, dh, dw), np.float32)
    with fm.Deconvolver(dh, dw, list(weights), [x.copy() for x in kernels], batch_planes=5, chunk_planes=2) as dec:
        dec.run_device([t.ptr for t in dviews], Z, 6, out_ptr=dout.ptr, est_ptr=dest.ptr,
                       stream=0)
        fm.device_synchronize()
    assert np.array_equal(from_device(dout, np.uint16), ref_out)
    assert np.array_equal(from_device(dest, np.float32), ref_est)


def test_two_lanes_match_one(fm, monkeypatch):
    """FMVD_LANES=2 runs alternate chunks on a second stream; chunks touch
    disjoint planes, so the results equal the single-stream run bit for bit
    (streaming and device-resident entry points)."""
    Z, V, dh, dw, k = 9, 3, 70, 96, 13
    data, kernels, weights = make_case(Z, V, dh, dw, k, k, 16, seed=17)
    monkeypatch.setenv("FMVD_LANES", "1")
    ref_out, ref_est, _ = run_gpu(fm, data, kernels, weights, 0, 5, 16, batch=9, chunk=2)
    monkeypatch.setenv("FMVD_LANES", "2")
    for b, c in ((9, 2), (9, 1), (5, 2)):
        out, est, _ = run_gpu(fm, data, kernels, weights, 0, 5, 16, batch=b, chunk=c)
        assert np.array_equal(out, ref_out) and np.array_equal(est, ref_est), (b, c)
    dviews = [to_device(np.ascontiguousarray(data[:, v])) for v in range(V)]
    dout = empty_device((Z, dh, dw), np.uint16)
    with fm.Deconvolver(dh, dw, list(weights), [x.copy() for x in kernels], batch_planes=9, chunk_planes=2) as dec:
        dec.run_device([t.ptr for t in dviews], Z, 5, out_ptr=dout.ptr, stream=0)
        fm.device_synchronize()
    assert np.array_equal(from_device(dout, np.uint16), ref_out)


def test_streaming_order_and_counts(fm):
    # planes with distinct constant levels must come back in z order
    Z, V, dh, dw = 5, 2, 40, 48
    data = np.stack([np.full((V, dh, dw), 100 * (z + 1), np.uint16) for z in range(Z)])
    k = np.ones((5, 5), np.float32)
    out, est, _ = run_gpu(fm, data, np.stack([k, k]), np.ones((V, dh, dw), np.float32), 0, 3, 16, batch=2)
    assert out.shape[0] == Z
    # uniform data c, two views of weight 1/2: E <- E + (c - E)/2 per view, E0 = 20
    for z in range(Z):
        c = 100.0 * (z + 1)
        e = c - (c - 20.0) * 0.5 ** 6
        assert np.abs(est[z] - e).max() < 1e-4 * c, z


@pytest.mark.parametrize("nplanes", [0, 1, 2])
def test_fewer_planes_than_streams(fm, nplanes):
    V, dh, dw = 2, 32, 32
    data = np.full((nplanes, V, dh, dw), 300, np.uint16)
    k = np.ones((3, 3), np.float32)
    out, est, _ = run_gpu(fm, data, np.stack([k, k]), np.ones((V, dh, dw), np.float32), 0, 2, 16, nstreams=3)
    assert out.shape == (nplanes, dh, dw)


@pytest.mark.parametrize("dh, views, want", [(1024, 4, 8), (2048, 4, 16), (200, 2, 256)])
def test_default_chunk_size(fm, dh, views, want):
    """Planes per chunk: the power of two whose per-iteration working set
    (spectrum and estimate per plane; kernel spectra and weights per view; the
    samples are streamed) fits the 200 MiB cache budget (8 at FFT 1536^2 with
    4 views); when the cache cannot hold the planes that fill the chip once,
    4 x that count (16 at FFT 2560^2); small planes are capped only by 4096
    and the batch."""
    k = np.ones((41, 41), np.float32)
    w = [np.ones((dh, dh), np.float32) for _ in range(views)]
    with fm.Deconvolver(dh, dh, w, [k.copy() for _ in range(views)], batch_planes=1024) as dec:
        assert dec.info()["chunk"] == want


def test_deterministic_rerun(fm):
    data, kernels, weights = make_case(2, 2, 120, 100, 15, 15, 16, seed=9)
    a = run_gpu(fm, data, kernels, weights, 1, 5, 16)
    b = run_gpu(fm, data, kernels, weights, 1, 5, 16)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_warm_start_matches_oracle(fm, orc):
    Z, V, dh, dw, k = 2, 2, 80, 96, 13
    data, kernels, weights = make_case(Z, V, dh, dw, k, k, 16, seed=11)
    rng = np.random.default_rng(1)
    init = rng.uniform(50, 400, size=(Z, dh, dw)).astype(np.float32)
    out_o, est_o = orc.rl_planes(data, weights, kernels.copy(), 0, 4, init=init.astype(np.float64))
    dviews = [to_device(np.ascontiguousarray(data[:, v])) for v in range(V)]
    dinit = to_device(init)
    dout = empty_device((Z, dh, dw), np.uint16)
    dest = empty_device((Z, dh, dw), np.float32)
    with fm.Deconvolver(dh, dw, list(weights), [x.copy() for x in kernels]) as dec:
        dec.run_device([t.ptr for t in dviews], Z, 4, out_ptr=dout.ptr, est_ptr=dest.ptr,
                       init_ptr=dinit.ptr, stream=0)
        fm.device_synchronize()
    check_rl(from_device(dout, np.uint16), from_device(dest, np.float32), out_o, est_o)
    # zero iterations return the warm start itself (through the pad/FFT round trip)
    dest0 = empty_device((Z, dh, dw), np.float32)
    with fm.Deconvolver(dh, dw, list(weights), [x.copy() for x in kernels]) as dec:
        dec.run_device([t.ptr for t in dviews], Z, 0, est_ptr=dest0.ptr, init_ptr=dinit.ptr,
                       stream=0)
        fm.device_synchronize()
    e0 = from_device(dest0, np.float32)
    assert np.abs(e0 - init).max() <= 1e-4 * np.abs(init).max()


def test_invalid_device_reports_error(fm):
    fm.lib().setCudaDevice(97)
    errs = fm.take_errors()
    assert errs and "HIP error" in errs[0]
    fm.lib().setCudaDevice(0)
    fm.check()
    # the reported error must not resurface in the next kernel launch check
    with fm.Deconvolver(16, 16, [np.ones((16, 16), np.float32)], [np.ones((3, 3), np.float32)]) as dec:
        out, _ = dec.run([[np.full((16, 16), 7, np.uint16)]], 1)
    fm.check()
    assert out.shape == (1, 16, 16)


def test_convergence_trace_matches_oracle_estimates(fm, orc):
    """errs[k] = ||E_k - ref|| / ||ref||, with ref the FP64 oracle's estimate
    after 4 iterations: errs[4] is the FP32/FP64 gap, and errs[k] for k < 4
    matches the distance between the oracle's own k- and 4-iteration estimates."""
    Z, V, dh, dw, k = 3, 2, 90, 110, 15
    data, kernels, weights = make_case(Z, V, dh, dw, k, k, 16, seed=21)
    ests = [orc.rl_planes(data, weights, kernels.copy(), 0, it, threads=8)[1] for it in range(5)]
    ref = ests[4].astype(np.float32)
    dviews = [to_device(np.ascontiguousarray(data[:, v])) for v in range(V)]
    dref = to_device(ref)
    with fm.Deconvolver(dh, dw, list(weights), [x.copy() for x in kernels], batch_planes=3, chunk_planes=2) as dec:
        errs = dec.convergence([t.ptr for t in dviews], Z, 4, dref.ptr)
    want = [np.linalg.norm(e - ests[4]) / np.linalg.norm(ests[4]) for e in ests]
    print("\\n[convergence]", errs, want)
    assert errs[4] < 1e-5
    for a, b in zip(errs[:4], want[:4]):
        assert abs(a - b) <= 1e-4 * b + 1e-6


def test_autotune_cache_file(fm, tmp_path):
    """With a cache file set, the first autotune of a configuration times the
    variants and records them; a second plan of the same configuration gets
    the same picks from the file without timing (one line, unchanged)."""
    path = tmp_path / "tune.txt"
    data, kernels, weights = make_case(1, 2, 600, 500, 41, 41, 16, seed=4)
    fm.lib().fmvd_set_autotune_cache(str(path).encode())
    try:
        with fm.Deconvolver(600, 500, list(weights), [k.copy() for k in kernels]) as dec:
            first = dec.autotune(planes=4)
        lines = path.read_text().splitlines()
        assert len(lines) == 1 and lines[0].split()[0].endswith("|p4")
        with fm.Deconvolver(600, 500, list(weights), [k.copy() for k in kernels]) as dec:
            second = dec.autotune(planes=4)
        assert {k: c for k, (c, _) in first.items()} == {k: c for k, (c, _) in second.items()}
        assert path.read_text().splitlines() == lines
    finally:
        fm.lib().fmvd_set_autotune_cache(None)
