"""Benchmark: full-size multi-view reconstruction on one MI355X.

Headline workload: four uint16 views of 2048 x 2048 x 1024 (80k Gaussian
beads on background 10, per-view anisotropic Gaussian PSF, Poisson noise),
each acquired through a random affine (rotation +-15 deg about each axis,
scale 0.9-1.1, translation +-20 voxels).  One reconstruction is the
GPU-resident chain fmvd_reconstruct_device_cuda_16: every view is resampled
back with the trilinear transform, its border mask becomes its weight image,
then RL runs with 41 x 41 PSFs (FFT extent 2560 x 2560) over all 1024 planes.
Inputs are resident in HBM when timing starts; data is synthetic and
generated on the GPU.

A step is one RL iteration over the whole volume (4 views x 1024 planes).
The warmup is one untimed reconstruction of `warmup` iterations; the timed
region is one reconstruction of exactly `steps` iterations, transforms and
mask weights included.  value = V * W * H * D * steps / elapsed.

The JSON line also carries:
* roofline: the dominant RL kernel's algorithmic bytes per launch (column
  pass 16 S, ratio pass 16 S + 2 P, update pass 16 S + 8 P per plane, with
  S = fftH (fftW/2 + 1), P = fftH fftW) over its mean launch time from HIP
  events attached to its dispatches;
* time_to_1e-3: time to reconstruction, plain and learned, until the
  estimate of a full-width plane is within 1e-3 relative L2 of the FP64
  oracle's 50-iteration RL of that plane;
* cpu_baseline: the FP64 oracle (and its FP32 twin) on a bounded sample of
  the same registered views, on this host;
* streamed: the 1024 x 1024 x 512 workload through the reference's
  plane-streaming entry point from host planes (PCIe included);
* batched: 48 uneven volumes in hipGraph replays;
* learned_vs_plain: the warm-start network trained on 128 x 128 x 64 samples.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "view-voxel RL iters/s + time to 1e-3 rel-L2 vs FP64 RL: learned vs plain"
HBM_PEAK_GBS = 8000.0
TOL = 1e-3
JUDGE_ITERS = 50


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=30, help="RL iterations in the timed reconstruction")
    p.add_argument("--warmup", type=int, default=2, help="RL iterations of the untimed reconstruction")
    p.add_argument("--planes", type=int, default=1024, help="volume depth (default: the workload's 1024)")
    p.add_argument("--chunk-planes", type=int, default=0,
                   help="planes per launch of the headline plan (0 = the plan's automatic choice)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-judge", action="store_true", help="skip the time-to-1e-3 legs")
    p.add_argument("--no-streamed", action="store_true", help="skip the 1024^2 x 512 streamed leg")
    p.add_argument("--no-batch", action="store_true", help="skip the batched uneven-volume leg")
    p.add_argument("--no-learned", action="store_true", help="skip the learned warm-start legs")
    p.add_argument("--train-samples", type=int, default=2000)
    p.add_argument("--epochs", type=int, default=3)
    p.add_argument("--learned-study", action="store_true",
                   help="only the warm-start convergence study: networks trained to a loss plateau on labels "
                        "of 20 and 50 plain-RL iterations, trajectories against the FP64 50-iteration judge")
    p.add_argument("--traffic-run", action="store_true",
                   help="short run for HBM counter collection: 64 planes, 2 iterations, no other legs, then "
                        "a 1 GiB device-to-device copy as the counters' calibration kernel")
    args = p.parse_args()
    if args.traffic_run:
        args.planes, args.steps, args.warmup = 64, 2, 0
        args.no_cpu = args.no_judge = args.no_streamed = args.no_batch = args.no_learned = True
    return args


def cpu_threads():
    """Threads for the oracle: the host's CPU share next to one GPU (16),
    or fewer when this process may run on fewer cores."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n)), n


def measured_traffic(fh, fw, kind, variant, planes_per_launch):
    """HBM bytes per launch of `kind` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per plane, calibrated against a plain copy), or
    None when it holds no entry for this extent and plan variant."""
    path = os.path.join(ROOT, "profiles", "traffic_rl_kernels.json")
    try:
        with open(path) as f:
            e = json.load(f)[f"{fh}x{fw}"][kind]
    except (OSError, KeyError, ValueError):
        return None
    if e.get("variant") != variant:
        return None
    return e["bytes_per_plane"] * planes_per_launch


# ------------------------------------------------------------------ data
def make_volume(fm, synth, shape, V, beads, seed, K=41, border=32):
    """Raw views -> acquired (warped by A_v); returns acquired DeviceArrays,
    the maps A_v (target -> source of the registration) and the kernels."""
    raw = synth.gpu_views(shape, V, beads=beads, seed=seed)
    rng = np.random.default_rng(seed + 7)
    A = [synth.random_affine(rng, shape) for _ in range(V)]
    acquired = []
    for v in range(V):
        d = fm.DeviceArray(shape, np.uint16)
        fm.transform_device(raw[v].ptr, shape, d.ptr, shape, synth.invert_affine(A[v]))
        acquired.append(d)
    fm.device_synchronize()
    for r in raw:
        r.free()
    kernels = [synth.gaussian_kernel_2d(*synth.view_sigma((1.5, 1.5, 4.0), v)[:2], K) for v in range(V)]
    return acquired, A, kernels


# ------------------------------------------------------------------ legs
def kernel_stats(timing, fh, fw, planes_per_launch):
    S, P = fh * (fw // 2 + 1), fh * fw
    per_plane = {"column": 16 * S, "ratio": 16 * S + 2 * P, "update": 16 * S + 8 * P}
    ks = {}
    for k, b in per_plane.items():
        ms, n = timing[k]
        if n:
            avg = ms / n / 1000.0
            ks[k] = dict(total_ms=ms, launches=n, avg_ms=1000 * avg, planes_per_launch=planes_per_launch,
                         algorithmic_bytes_per_launch=b * planes_per_launch,
                         achieved_gbs=b * planes_per_launch / avg / 1e9)
    B_RL = 64 * S + 10 * P
    rl_ms = sum(ks[k]["total_ms"] for k in ks)
    iter_gbs = B_RL * ks["ratio"]["launches"] * planes_per_launch / (rl_ms / 1000.0) / 1e9
    return ks, B_RL, iter_gbs


def cpu_baseline(oracle, registered, weights, kernels, shape, planes=4, iters=3):
    """FP64 oracle and its FP32 twin on `planes` planes x all views x `iters`
    iterations of the registered volume; plan setup (a 0-iteration run) is
    subtracted."""
    Z, H, W = shape
    V = len(registered)
    z0 = Z // 2 - planes // 2
    data = np.ascontiguousarray(np.stack([r.download_planes(z0, planes) for r in registered], axis=1))
    kern, wts = np.stack(kernels), np.stack(weights)
    nt, ncpu = cpu_threads()
    out = {}
    for prec in ("f64", "f32"):
        # the first call also pays for first-touch page faults of its buffers
        oracle.rl_planes(data, wts, kern.copy(), 0, 0, threads=nt, precision=prec)
        t = time.perf_counter()
        oracle.rl_planes(data, wts, kern.copy(), 0, 0, threads=nt, precision=prec)
        t_setup = time.perf_counter() - t
        t = time.perf_counter()
        oracle.rl_planes(data, wts, kern.copy(), 0, iters, threads=nt, precision=prec)
        t_run = max(time.perf_counter() - t - t_setup, 1e-9)
        out[prec] = dict(value=V * W * H * planes * iters / t_run, seconds=t_run, setup_s=t_setup)
        log(f"[bench] cpu {prec}: {out[prec]['value']:.4g} view-voxel it/s on {nt} threads ({t_run:.1f}s)")
    return {"value": out["f64"]["value"], "unit": "view-voxel RL iterations/s", "cores": nt, "kind": "port",
            "sample": f"FP64 oracle (the FP32 twin in 'fp32'), {planes} planes x {V} views x {iters} iteration of "
                      f"the registered full-size volume (FFT {snap(H + 40)}x{snap(W + 40)}), OpenMP inside each "
                      f"plane's transforms on {nt} of the {ncpu} cores this process may use; plan setup subtracted",
            "fp32": out["f32"]}


def snap(n):
    import gpu_deconvolution_amd as fm
    return fm.snap_transform_size(n)


def judge_plane(oracle, registered, weights, kernels, z, iters=JUDGE_ITERS):
    """The FP64 oracle's `iters`-iteration cold-start RL of full-width plane z."""
    data = np.ascontiguousarray(np.stack([r.download_planes(z, 1) for r in registered], axis=1))
    nt, _ = cpu_threads()
    t = time.perf_counter()
    _, ref = oracle.rl_planes(data, np.stack(weights), np.stack(kernels).copy(), 0, iters, threads=nt)
    return ref[0], time.perf_counter() - t


def first_below(errs, tol=TOL):
    return next((k for k, e in enumerate(errs) if e <= tol), None)


def streamed_leg(fm, synth, seed=1001, iters=30):
    """The 1024 x 1024 x 512 workload (20k beads, random affines, 41 x 41
    PSFs, FFT 1536^2) through fmvd_deconvolve_planes_cuda_16: registered
    planes streamed from host memory through the datasource and results
    returned through the datasink (plane-streamed double buffering, PCIe
    included)."""
    Z, H, W, V = 512, 1024, 1024, 4
    shape = (Z, H, W)
    acquired, A, kernels = make_volume(fm, synth, shape, V, 20000, seed)
    host, weights = [], []
    reg = fm.DeviceArray(shape, np.uint16)
    mask = fm.DeviceArray((H, W), np.float32)
    for v in range(V):
        fm.transform_device(acquired[v].ptr, shape, reg.ptr, shape, A[v])
        fm.mask_device(shape, shape, A[v], 32, 1.0, mask.ptr)
        fm.device_synchronize()
        host.append(reg.download())
        weights.append(mask.download())
        acquired[v].free()
    reg.free()
    mask.free()
    planes = [[host[v][z] for v in range(V)] for z in range(Z)]
    with fm.Deconvolver(H, W, weights, [k.copy() for k in kernels], psf_type=0, bits=16) as dec:
        info = dec.info()
        dec.autotune(planes=info["chunk"])
        dec.run(planes[:info["chunk"]], 1)  # warm-up
        t = time.perf_counter()
        out, _ = dec.run(planes, iters)
        el = time.perf_counter() - t
    assert out.shape == shape
    rate = V * Z * H * W * iters / el
    log(f"[bench] streamed 1024^2 x 512: {iters} iterations in {el:.2f}s -> {rate:.4g} view-voxel it/s "
        f"(batch {info['batch']}, chunk {info['chunk']})")
    return {"workload": f"4 views 1024x1024x512 uint16, 20k beads, random affines, 41x41 PSFs, {iters} iterations "
                        f"through fmvd_deconvolve_planes_cuda_16 from host planes (FFT {info['fft_h']}x{info['fft_w']})",
            "value": rate, "unit": "view-voxel RL iterations/s", "seconds": el, "pcie_inclusive": True,
            "batch_planes": info["batch"], "planes_per_launch": info["chunk"]}


def batched_workload(fm, synth, seed=2024, nvol=48, iters=25, kernel=31):
    """Batched RL over 48 independent reconstructions of uneven size (xy in
    {128, 192, 256, 384, 512}, z in 32..256, 2-4 views, per-volume PSF sigma
    xy 1-2.5 / z 3-6, Poisson uint16), 25 iterations replayed as a hipGraph,
    FFT plans chosen by the batch autotuner."""
    rng = np.random.default_rng(seed)
    vols, vvi, brl = [], 0, 0
    t0 = time.time()
    for i in range(nvol):
        xy = int(rng.choice([128, 192, 256, 384, 512]))
        z = int(rng.integers(32, 257))
        V = int(rng.integers(2, 5))
        sxy, sz = float(rng.uniform(1.0, 2.5)), float(rng.uniform(3.0, 6.0))
        views = synth.gpu_views((z, xy, xy), V, beads=max(50, z * xy * xy // 20000), seed=seed + i,
                                sigma=(sxy, sxy, sz))
        kernels = [synth.gaussian_kernel_2d(*synth.view_sigma((sxy, sxy, sz), v)[:2], kernel) for v in range(V)]
        vols.append(dict(views=views, weights=[np.ones((xy, xy), np.float32) for _ in range(V)], kernels=kernels,
                         out=fm.DeviceArray((z, xy, xy), np.uint16)))
        vvi += V * z * xy * xy
        fh = fw = fm.snap_transform_size(xy + kernel - 1)
        brl += V * z * (64 * fh * (fw // 2 + 1) + 10 * fh * fw)  # B_RL of every (view, plane), one iteration
    log(f"[bench] batched workload: {nvol} volumes generated in {time.time() - t0:.1f}s")
    b = fm.Batch(vols)
    classes = b.info()
    tuned = b.autotune()
    b.run(1, use_graph=True)  # warm-up (captures the graph)
    fm.device_synchronize()
    t = time.perf_counter()
    b.run(iters, use_graph=True)
    fm.device_synchronize()
    el = time.perf_counter() - t
    log(f"[bench] batched: {iters} iterations in {el:.3f}s -> {vvi * iters / el:.4g} view-voxel it/s, "
        f"{brl * iters / el / 1e9:.0f} GB/s algorithmic; "
        f"classes {classes}, variants {tuned}")
    b.close()
    for v in vols:
        for d in v["views"]:
            d.free()
        v["out"].free()
    return {"workload": f"{nvol} independent reconstructions, xy in {{128,192,256,384,512}}, z 32-256, 2-4 views, "
                        f"{kernel}x{kernel} PSFs, {iters} iterations as hipGraph replays",
            "value": vvi * iters / el, "unit": "view-voxel RL iterations/s", "seconds": el,
            "classes": classes, "fft_plan_variants": tuned,
            "rl_iteration_roofline": {"achieved": brl * iters / el / 1e9, "frac": brl * iters / el / 1e9 / HBM_PEAK_GBS,
                                      "unit": "GB/s", "note": "B_RL = 64 S + 10 P per view, iteration and plane of "
                                      "each volume's FFT class; the 256^2 and 512^2 classes are cache-resident, "
                                      "so the figure can exceed what HBM alone allows"}}


def learned_leg(fm, oracle, ntrain=2000, epochs=3, label_iters=50, ntest=2, lr=2e-4, seed=7, tol=TOL):
    """Learned warm start vs plain RL on the 128 x 128 x 64, 2-view training
    workload: train the network on `ntrain` (views -> `label_iters`-iteration
    plain RL) pairs, then on held-out seeds compare time to a `tol`
    relative-L2 stop against the FP64 oracle's `label_iters`-iteration RL
    (the judge).  The warm start is rejected (plain RL is the fallback) when
    its trajectory does not reach the stop within label_iters + 20
    iterations.  Returns (the JSON dict, the trained network)."""
    from gpu_deconvolution_amd import warmstart as ws
    shape, V = (64, 128, 128), 2
    Z, H, W = shape
    t0 = time.time()
    samples = ws.make_samples(ntrain + ntest, shape, V, 31, label_iters, seed=seed, log=log)
    train, test = samples[:ntrain], samples[ntrain:]
    gen_s = time.time() - t0
    net = fm.Net(seed)
    tr = ws.Trainer(net, shape, V, batch=8)
    t0 = time.time()
    hist = tr.fit(train, epochs, lr=lr, log=log, lr_decay=0.75)
    fm.device_synchronize()
    train_s = time.time() - t0
    tr.free()
    for s in train:
        s.free()
    results = []
    nt, _ = cpu_threads()
    for s in test:
        data = np.stack([v.download() for v in s.views], 1)
        t = time.perf_counter()
        _, ref = oracle.rl_planes(data, np.stack(s.weights), np.stack(s.kernels).copy(), 0, label_iters, threads=nt)
        oracle_s = time.perf_counter() - t
        dref = fm.DeviceArray.from_host(ref.astype(np.float32))
        init = fm.DeviceArray(shape, np.float32)
        ptrs = [v.ptr for v in s.views]
        with fm.Deconvolver(H, W, s.weights, [k.copy() for k in s.kernels]) as dec:
            ep = dec.convergence(ptrs, Z, label_iters + 20, dref.ptr)
            fm.device_synchronize()
            t = time.perf_counter()
            ws.warm_start(net, s.views, s.scale, init)
            fm.device_synchronize()
            infer_s = time.perf_counter() - t
            ew = dec.convergence(ptrs, Z, label_iters + 20, dref.ptr, init_ptr=init.ptr)
            kp, kw = first_below(ep, tol), first_below(ew, tol)

            def timed(iters, init_ptr=0):
                fm.device_synchronize()
                t1 = time.perf_counter()
                dec.run_device(ptrs, Z, iters, init_ptr=init_ptr)
                fm.device_synchronize()
                return time.perf_counter() - t1

            timed(2)
            plain_s = timed(kp) if kp is not None else None
            warm_iter_s = timed(kw, init.ptr) if kw is not None else None
        best = float(ew.min())
        k_equiv = next((k for k, e in enumerate(ep) if e <= best), None)
        accepted = kw is not None and plain_s is not None and infer_s + warm_iter_s < plain_s
        learned_s = infer_s + warm_iter_s if accepted else infer_s + (plain_s or 0.0)
        saved, saved_j, saved_e = iterations_saved(ep, ew)
        results.append(dict(k_plain=kp, k_warm=kw, accepted=accepted, plain_s=plain_s, learned_s=learned_s,
                            max_iterations_saved=saved, saved_at_warm_k=saved_j, saved_at_error=saved_e,
                            infer_s=infer_s, speedup=(plain_s / learned_s) if plain_s and learned_s else None,
                            net_rel_l2=float(ew[0]), best_warm_rel_l2=best, best_warm_k=int(ew.argmin()),
                            plain_k_to_best=k_equiv, oracle_s=oracle_s,
                            plain_errs=[float(x) for x in ep[::5]], warm_errs=[float(x) for x in ew[::5]]))
        log(f"[bench] learned: k_plain={kp} k_warm={kw} accepted={accepted} net relL2 {ew[0]:.3e} "
            f"best warm {best:.3e}@{int(ew.argmin())} (plain reaches it at {k_equiv}); "
            f"plain {plain_s} s, learned {learned_s:.4f} s")
        for b in (dref, init):
            b.free()
        s.free()
    leg = {"workload": f"learned warm start, {ntrain} training samples 128x128x64 x 2 views (labels: "
                       f"{label_iters}-iteration plain RL), {epochs} epochs batch 8 Adam; held-out seeds run to a "
                       f"{tol:g} relative-L2 stop vs the FP64 oracle's {label_iters}-iteration RL",
           "generate_s": gen_s, "train_s": train_s, "loss_per_epoch": hist, "samples": results,
           "speedup": float(np.mean([r["speedup"] for r in results])) if results else None}
    return leg, net


def iterations_saved(plain, warm):
    """For every point j of the warm trajectory: the plain iterations needed
    to reach the same error, minus j.  Returns the best (saved, j, error)."""
    best = (0, 0, float(warm[0]))
    for j, e in enumerate(warm):
        k = next((k for k, p in enumerate(plain) if p <= e), None)
        if k is not None and k - j > best[0]:
            best = (k - j, j, float(e))
    return best


def learned_study(fm, oracle, ntrain=2000, epochs=10, label_sets=(20, 50), judge_iters=50, ntest=4, seed=11):
    """Convergence study of the warm start: for each label depth k0, a network
    trained until its loss stops falling on (views -> k0-iteration plain RL)
    pairs; on held-out seeds, the relative-L2 distance to the FP64 oracle's
    judge_iters-iteration RL after every iteration of plain RL and of RL
    started from the network."""
    from gpu_deconvolution_amd import warmstart as ws
    shape, V = (64, 128, 128), 2
    Z, H, W = shape
    nt, _ = cpu_threads()
    out = {"workload": f"{ntrain} training samples 128x128x64 x 2 views, batch 8 Adam, up to {epochs} epochs "
                       f"(stop when the epoch loss falls by less than 2%); {ntest} held-out seeds; judge: FP64 "
                       f"oracle {judge_iters}-iteration cold-start RL", "nets": []}
    test = ws.make_samples(ntest, shape, V, 31, judge_iters, seed=seed + 999, log=log)
    refs = []
    for s in test:
        data = np.stack([v.download() for v in s.views], 1)
        _, ref = oracle.rl_planes(data, np.stack(s.weights), np.stack(s.kernels).copy(), 0, judge_iters, threads=nt)
        refs.append(fm.DeviceArray.from_host(ref.astype(np.float32)))
    plain = []
    for s, dref in zip(test, refs):
        with fm.Deconvolver(H, W, s.weights, [k.copy() for k in s.kernels]) as dec:
            plain.append(dec.convergence([v.ptr for v in s.views], Z, judge_iters + 20, dref.ptr))
    for k0 in label_sets:
        samples = ws.make_samples(ntrain, shape, V, 31, k0, seed=seed, log=log)
        net = fm.Net(seed)
        tr = ws.Trainer(net, shape, V, batch=8)
        hist, lr = [], 4e-4
        for ep in range(epochs):
            hist += tr.fit(samples, 1, lr=lr, seed=ep, log=log)
            lr *= 0.8
            if len(hist) >= 3 and hist[-1] > 0.98 * hist[-2]:
                break
        tr.free()
        for s in samples:
            s.free()
        rows = []
        init = fm.DeviceArray(shape, np.float32)
        for s, dref, ep in zip(test, refs, plain):
            ws.warm_start(net, s.views, s.scale, init)
            with fm.Deconvolver(H, W, s.weights, [k.copy() for k in s.kernels]) as dec:
                ew = dec.convergence([v.ptr for v in s.views], Z, judge_iters + 20, dref.ptr, init_ptr=init.ptr)
            saved, j, e = iterations_saved(ep, ew)
            rows.append(dict(net_rel_l2=float(ew[0]), best_warm_rel_l2=float(ew.min()), best_warm_k=int(ew.argmin()),
                             k_warm=first_below(ew), k_plain=first_below(ep), max_saved=saved, saved_at_k=j,
                             saved_at_error=e, warm_errs=[float(x) for x in ew[:41:2]],
                             plain_errs=[float(x) for x in ep[:41:2]]))
            log(f"[study] k0={k0}: net {ew[0]:.3e}, best warm {ew.min():.3e}@{int(ew.argmin())}, saves {saved} "
                f"iterations at error {e:.3e} (warm iteration {j}); plain reaches 1e-3 at {first_below(ep)}, "
                f"warm at {first_below(ew)}")
        init.free()
        net.close()
        out["nets"].append(dict(label_iterations=k0, loss_per_epoch=hist, samples=rows))
    for r in refs:
        r.free()
    for s in test:
        s.free()
    return out


# ------------------------------------------------------------------ main
def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import gpu_deconvolution_amd as fm
    from gpu_deconvolution_amd import synth

    if args.learned_study:
        import oracle
        fm.lib().setCudaDevice(0)
        print(json.dumps({"learned_study": learned_study(fm, oracle)}), flush=True)
        return

    Z, H, W, V, K = args.planes, 2048, 2048, 4, 41
    shape = (Z, H, W)
    fm.lib().setCudaDevice(0)
    fm.check()
    name = C.create_string_buffer(256)
    fm.lib().getCudaDeviceName(0, name)

    t0 = time.time()
    acquired, A, kernels = make_volume(fm, synth, shape, V, int(round(80000 * Z / 1024)), seed=3003, K=K)
    registered = [fm.DeviceArray(shape, np.uint16) for _ in range(V)]
    out = fm.DeviceArray(shape, np.uint16)
    log(f"[bench] data ready in {time.time() - t0:.1f}s on {name.value.decode()}")

    dec = fm.Deconvolver(H, W, None, [k.copy() for k in kernels], psf_type=0, bits=16, batch_planes=min(Z, 256),
                         chunk_planes=args.chunk_planes)
    info = dec.info()
    lanes = int(os.environ.get("FMVD_LANES", "1") or 1)
    stream = fm.Stream()

    def reconstruct(iters):
        dec.reconstruct_device([a.ptr for a in acquired], shape, A, 32, 1.0, [r.ptr for r in registered], Z,
                               iters, out_ptr=out.ptr, stream=stream.handle)
        stream.synchronize()

    reconstruct(0)  # registered volumes and mask weights for the autotuner
    t0 = time.time()
    tuned = dec.autotune(planes=info["chunk"])
    variants = {k: c for k, (c, _) in tuned.items()}
    log(f"[bench] plan {info}; autotune in {time.time() - t0:.1f}s: " +
        ", ".join(f"{k}=v{c} {['%.3f' % x if x is not None else '-' for x in ts]}" for k, (c, ts) in tuned.items()))

    t = time.time()
    reconstruct(args.warmup)
    log(f"[bench] warmup reconstruction ({args.warmup} iterations): {time.time() - t:.3f}s")
    fm.device_synchronize()
    with fm.KernelTiming() as kt:
        t = time.perf_counter()
        reconstruct(args.steps)
        elapsed = time.perf_counter() - t
        timing = {k: kt.read(k) for k in ("column", "ratio", "update", "ingest", "egress", "transform")}
    ms_per_step = 1000.0 * elapsed / args.steps
    value = V * W * H * Z * args.steps / elapsed
    log(f"[bench] {args.steps} iterations (+4 transforms, mask weights) in {elapsed:.3f}s -> {ms_per_step:.1f} "
        f"ms/step, {value:.4g} view-voxel it/s (chunk {info['chunk']}, lanes {lanes})")

    fh, fw = info["fft_h"], info["fft_w"]
    planes_per_launch = min(Z, info["batch"], info["chunk"])
    assert Z % planes_per_launch == 0, "timed chunks must be full"
    kstats, B_RL, iter_gbs = kernel_stats(timing, fh, fw, planes_per_launch)
    dom = max(kstats, key=lambda k: kstats[k]["total_ms"])
    tr_ms, tr_n = timing["transform"]
    tr_gbs = (2 * Z * H * W * 2) * tr_n / (tr_ms / 1000.0) / 1e9 if tr_n else None
    for k, s in kstats.items():
        log(f"[bench] {k:7s} {s['launches']:5d} timed launches of {s['planes_per_launch']} planes, avg "
            f"{1000 * s['avg_ms']:.1f} us  {s['achieved_gbs']:.0f} GB/s")
    log(f"[bench] RL chain (B_RL = 64S+10P = {B_RL / 1e6:.2f} MB/view-iter-plane): {iter_gbs:.0f} GB/s; "
        f"transform {tr_gbs and round(tr_gbs)} GB/s ({tr_n and tr_ms / tr_n:.2f} ms per view)")

    # masks as the plan holds them (normalised on the device from these)
    weights = []
    mask = fm.DeviceArray((H, W), np.float32)
    for v in range(V):
        fm.mask_device(shape, shape, A[v], 32, 1.0, mask.ptr)
        weights.append(mask.download())
    mask.free()

    cpu = None
    if not args.no_cpu:
        import oracle
        cpu = cpu_baseline(oracle, registered, weights, kernels, shape)

    tt = None
    if not args.no_judge:
        import oracle
        zj = Z // 2
        ref, oracle_s = judge_plane(oracle, registered, weights, kernels, zj)
        log(f"[bench] judge: FP64 oracle {JUDGE_ITERS} iterations of plane {zj} in {oracle_s:.1f}s")
        dref = fm.DeviceArray.from_host(ref.astype(np.float32))
        pv = [r.ptr + zj * H * W * 2 for r in registered]
        ep = dec.convergence(pv, 1, JUDGE_ITERS + 10, dref.ptr)
        kp = first_below(ep)
        plain_s = None
        if kp is not None:
            t = time.perf_counter()
            reconstruct(kp)
            plain_s = time.perf_counter() - t
        log(f"[bench] time to {TOL:g}: plain needs {kp} iterations, {plain_s} s (errors {ep[::10]})")
        tt = {"judge": f"FP64 oracle, {JUDGE_ITERS}-iteration cold-start RL of full-width plane z={zj} (the "
                       f"512x512x256 crop's centre plane), relative L2 over the plane's data region",
              "oracle_s": oracle_s, "k_plain": kp, "plain_s": plain_s,
              "plain_errs": [float(x) for x in ep[::5]]}
        keep_ref = (dref, pv)
    streamed = batched = None
    learned = net = None
    if not args.no_learned and not args.no_cpu:
        import oracle
        learned, net = learned_leg(fm, oracle, ntrain=args.train_samples, epochs=args.epochs)
    if tt is not None:
        dref, pv = keep_ref
        if net is not None:
            from gpu_deconvolution_amd import warmstart as ws
            init = fm.DeviceArray(shape, np.float32)
            fm.device_synchronize()
            t = time.perf_counter()
            # per-volume scale: the mean sample over 16 central planes of every view
            scale = max(1.0, float(np.mean([r.download_planes(Z // 2 - 8, 16).mean(dtype=np.float64)
                                            for r in registered])))
            net.infer(registered, scale, init, floor_value=1e-3 * scale, tile=(128, 256, 256, 16))
            infer_s = time.perf_counter() - t
            ew = dec.convergence(pv, 1, JUDGE_ITERS + 10, dref.ptr, init_ptr=init.ptr + zj * H * W * 4)
            kw = first_below(ew)
            warm_s = None
            if kw is not None:
                t = time.perf_counter()
                for a, r, m in zip(acquired, registered, A):
                    fm.transform_device(a.ptr, shape, r.ptr, shape, m, stream=stream.handle)
                dec.run_device([r.ptr for r in registered], Z, kw, out_ptr=out.ptr, init_ptr=init.ptr,
                               stream=stream.handle)
                stream.synchronize()
                warm_s = time.perf_counter() - t + infer_s
            accepted = warm_s is not None and tt["plain_s"] is not None and warm_s < tt["plain_s"]
            learned_s = warm_s if accepted else infer_s + (tt["plain_s"] or 0.0)
            tt.update({"infer_s": infer_s, "tile": "256x256x128 interior tiles, 16-voxel halo", "k_warm": kw,
                       "net_rel_l2": float(ew[0]), "best_warm_rel_l2": float(ew.min()),
                       "best_warm_k": int(ew.argmin()), "accepted": accepted, "learned_s": learned_s,
                       "speedup": (tt["plain_s"] / learned_s) if tt["plain_s"] and learned_s else None,
                       "warm_errs": [float(x) for x in ew[::5]]})
            log(f"[bench] learned at full size: inference {infer_s:.2f}s, warm start relL2 {ew[0]:.3e}, "
                f"k_warm={kw}, accepted={accepted}, learned {learned_s} s vs plain {tt['plain_s']} s")
            init.free()
        dref.free()

    for r in registered + acquired + [out]:
        r.free()
    dec.close()
    stream.close()
    if not args.no_streamed:
        streamed = streamed_leg(fm, synth)
    if not args.no_batch:
        batched = batched_workload(fm, synth)
    if net is not None:
        net.close()

    dk = kstats[dom]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "view-voxel RL iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (GPU-generated beads, Poisson noise, random affines)",
        "config": {
            "workload": f"full-size reconstruction: 4 views uint16 2048x2048x{Z}, 80k beads, random affines, "
                        f"GPU transform back + border-mask weights + plain RL, 41x41 PSFs, FFT {fh}x{fw}; "
                        f"a step is one RL iteration over the whole volume",
            "views": V, "shape_zyx": [Z, H, W], "kernel": K, "iterations_timed": args.steps, "fft": [fh, fw],
            "planes_per_launch": planes_per_launch, "device": name.value.decode(), "fft_plan_variants": variants,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": dk["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dk["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": measured_traffic(fh, fw, dom, variants[dom], planes_per_launch),
        },
        "rl_iteration_roofline": {"achieved": iter_gbs, "frac": iter_gbs / HBM_PEAK_GBS, "unit": "GB/s",
                                  "bytes_per_view_iter_plane": B_RL},
        "kernels": kstats,
        "transform_gbs": tr_gbs,
        "cpu_baseline": cpu,
        "time_to_1e-3": tt,
        "streamed": streamed,
        "batched": batched,
        "learned_vs_plain": learned,
    }
    print(json.dumps(line), flush=True)

    if args.traffic_run:
        # calibration kernel for FETCH_SIZE / WRITE_SIZE: a plain 1 GiB copy
        n = 1 << 30
        a, b = fm.DeviceArray((n,), np.uint8), fm.DeviceArray((n,), np.uint8)
        a.fill_zero()
        fm.device_synchronize()
        if synth.synth_lib().fmvd_synth_copy(a.ptr, b.ptr, n, None) != 0:
            fm.check()
        fm.device_synchronize()
        log(f"[bench] calibration copy of {n} bytes done")


if __name__ == "__main__":
    main()
