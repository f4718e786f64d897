"""Benchmark: affine transform + plane-wise multi-view RL at real size, one MI355X.

Workload (the metric's real-size configuration): four uint16 views of
1024 x 1024 x 512 (20k Gaussian beads on background 10, per-view anisotropic
Gaussian PSF, Poisson noise), each warped by a random affine (rotation
+-15 deg about each axis, scale 0.9-1.1, translation +-20 voxels).  One step
= resample the four views back with the trilinear transform and run 30 RL
iterations over all 512 planes with 41 x 41 PSFs (FFT extent 1536 x 1536),
weights = the transform's border masks.  Inputs are resident in HBM when the
timed region starts; data is synthetic and generated on the GPU.

Prints one JSON line.  value = view-voxel RL iterations per second over the
whole step (V * W * H * D * iterations / step time).  roofline: the
dominant kernel's algorithmic bytes per launch (from the per-plane pass
table: column pass 16 S, ratio pass 16 S + 2 P, update pass 16 S + 8 P bytes
with S = fftH (fftW/2 + 1), P = fftH fftW) over its mean launch time measured
with HIP events on its launch stream.  cpu_baseline: the FP64 oracle on a
bounded z-subset of the same registered views, timed on this host.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--planes", type=int, default=512, help="volume depth (default: the workload's 512)")
    p.add_argument("--iterations", type=int, default=30)
    p.add_argument("--cpu-planes", type=int, default=16)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-batch", action="store_true", help="skip the batched uneven-volume measurement")
    p.add_argument("--no-learned", action="store_true", help="skip the learned warm-start comparison")
    p.add_argument("--no-full", action="store_true", help="skip the full-size (2048x2048x1024) plain pipeline")
    p.add_argument("--train-samples", type=int, default=2000)
    p.add_argument("--epochs", type=int, default=3)
    p.add_argument("--traffic-run", action="store_true",
                   help="short run for HBM counter collection: 64 planes, 3 iterations, 1 step, no CPU "
                        "baseline, then a 1 GiB device-to-device copy as the counters' calibration kernel")
    args = p.parse_args()
    if args.traffic_run:
        args.planes, args.iterations, args.steps, args.warmup, args.no_cpu, args.no_batch = 64, 3, 1, 0, True, True
        args.no_learned = args.no_full = True
    return args


def measured_traffic(fh, fw, kind, planes_per_launch):
    """HBM bytes per launch of `kind` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per plane, calibrated against a plain copy), or
    None when no summary exists for this FFT extent."""
    path = os.path.join(ROOT, "profiles", "traffic_rl_kernels.json")
    try:
        with open(path) as f:
            table = json.load(f)
        per_plane = table[f"{fh}x{fw}"][kind]["bytes_per_plane"]
    except (OSError, KeyError, ValueError):
        return None
    return per_plane * planes_per_launch


def batched_workload(fm, synth, seed=2024, nvol=48, iters=25, kernel=31):
    """Batched RL over 48 independent reconstructions of uneven size (xy in
    {128, 192, 256, 384, 512}, z in 32..256, 2-4 views, per-volume PSF sigma
    xy 1-2.5 / z 3-6, Poisson uint16), 25 iterations replayed as a hipGraph,
    FFT plans chosen by the batch autotuner.  Returns a dict for the JSON line."""
    rng = np.random.default_rng(seed)
    vols, vvi = [], 0
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
    log(f"[bench] batched: {iters} iterations in {el:.3f}s -> {vvi * iters / el:.4g} view-voxel it/s; "
        f"classes {classes}, variants {tuned}")
    b.close()
    for v in vols:
        for d in v["views"]:
            d.free()
        v["out"].free()
    return {"workload": f"{nvol} independent reconstructions, xy in {{128,192,256,384,512}}, z 32-256, 2-4 views, "
                        f"{kernel}x{kernel} PSFs, {iters} iterations as hipGraph replays",
            "value": vvi * iters / el, "unit": "view-voxel RL iterations/s", "seconds": el,
            "classes": classes, "fft_plan_variants": tuned}


def full_size_leg(fm, synth, seed=3003, iters=30, kernel=41):
    """The full-size plain pipeline: 4 uint16 views of 2048 x 2048 x 1024
    (80k beads, random affines as in the real-size workload), transformed back
    on the GPU, then `iters` RL iterations with 41 x 41 PSFs (FFT 2560^2).
    Reports the RL rate and the time to reconstruction (transforms + RL) of
    one timed run after a one-iteration warm-up.  Returns a dict for the JSON
    line."""
    Z, H, W, V = 1024, 2048, 2048, 4
    shape = (Z, H, W)
    t0 = time.time()
    raw = synth.gpu_views(shape, V, beads=80000, seed=seed)
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
    kernels = [synth.gaussian_kernel_2d(*synth.view_sigma((1.5, 1.5, 4.0), v)[:2], kernel) for v in range(V)]
    weights = []
    mask = fm.DeviceArray((H, W), np.float32)
    for v in range(V):
        fm.mask_device(shape, shape, A[v], 32, 1.0, mask.ptr)
        weights.append(mask.download())
    mask.free()
    registered = [fm.DeviceArray(shape, np.uint16) for _ in range(V)]
    out = fm.DeviceArray(shape, np.uint16)
    log(f"[bench] full size: data ready in {time.time() - t0:.1f}s")
    dec = fm.Deconvolver(H, W, weights, [k.copy() for k in kernels], psf_type=0, bits=16, batch_planes=256)
    info = dec.info()
    dec.autotune(planes=info["chunk"])
    stream = fm.Stream()

    def reconstruct(n_iter):
        for v in range(V):
            fm.transform_device(acquired[v].ptr, shape, registered[v].ptr, shape, A[v], stream=stream.handle)
        dec.run_device([r.ptr for r in registered], Z, n_iter, out_ptr=out.ptr, stream=stream.handle)
        stream.synchronize()

    reconstruct(1)
    t = time.perf_counter()
    reconstruct(iters)
    el = time.perf_counter() - t
    rate = V * Z * H * W * iters / el
    log(f"[bench] full size: {iters} iterations + 4 transforms in {el:.2f}s -> {rate:.4g} view-voxel it/s "
        f"(plan {info})")
    dec.close()
    for d in acquired + registered + [out]:
        d.free()
    return {"workload": f"4 views of 2048x2048x1024 uint16, 80k beads, random affines, {kernel}x{kernel} PSFs, "
                        f"transform back + {iters} plain RL iterations (FFT {info['fft_h']}x{info['fft_w']})",
            "value": rate, "unit": "view-voxel RL iterations/s", "seconds_to_reconstruction": el,
            "planes_per_launch": info["chunk"]}


def learned_leg(fm, oracle, ntrain=2000, epochs=3, label_iters=50, ntest=2, lr=2e-4, seed=7, tol=1e-3):
    """Learned warm start vs plain RL (the 128 x 128 x 64, 2-view training
    workload): train the network on `ntrain` (views -> `label_iters`-iteration
    plain RL) pairs, then on held-out seeds compare time to a `tol` relative-L2
    stop against the FP64 oracle's `label_iters`-iteration RL (the judge).  The
    warm start is rejected (plain RL is the fallback) when its trajectory does
    not reach the stop within label_iters + 20 iterations."""
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
    for s in test:
        data = np.stack([v.download() for v in s.views], 1)
        t = time.perf_counter()
        _, ref = oracle.rl_planes(data, np.stack(s.weights), np.stack(s.kernels).copy(), 0, label_iters,
                                  threads=cpu_threads())
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
            kp = next((k for k, e in enumerate(ep) if e <= tol), None)
            kw = next((k for k, e in enumerate(ew) if e <= tol), None)

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
        results.append(dict(k_plain=kp, k_warm=kw, accepted=accepted, plain_s=plain_s, learned_s=learned_s,
                            infer_s=infer_s, speedup=(plain_s / learned_s) if plain_s and learned_s else None,
                            net_rel_l2=float(ew[0]), best_warm_rel_l2=best, best_warm_k=int(ew.argmin()),
                            plain_k_to_best=k_equiv, oracle_s=oracle_s))
        log(f"[bench] learned: k_plain={kp} k_warm={kw} accepted={accepted} net relL2 {ew[0]:.3e} "
            f"best warm {best:.3e}@{int(ew.argmin())} (plain reaches it at {k_equiv}); "
            f"plain {plain_s} s, learned {learned_s:.4f} s")
        for b in (dref, init):
            b.free()
        s.free()
    return {"workload": f"learned warm start, {ntrain} training samples 128x128x64 x 2 views (labels: "
                        f"{label_iters}-iteration plain RL), {epochs} epochs batch 8 Adam; held-out seeds run to a "
                        f"{tol:g} relative-L2 stop vs the FP64 oracle's {label_iters}-iteration RL",
            "generate_s": gen_s, "train_s": train_s, "loss_per_epoch": hist, "samples": results,
            "speedup": float(np.mean([r["speedup"] for r in results])) if results else None}


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import gpu_deconvolution_amd as fm
    from gpu_deconvolution_amd import synth

    Z, H, W, V, K, ITERS = args.planes, 1024, 1024, 4, 41, args.iterations
    seed = 1001
    shape = (Z, H, W)
    fm.lib().setCudaDevice(0)
    fm.check()
    name = C.create_string_buffer(256)
    fm.lib().getCudaDeviceName(0, name)

    t0 = time.time()
    raw = synth.gpu_views(shape, V, beads=20000, seed=seed)
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
    weights = []
    mask = fm.DeviceArray((H, W), np.float32)
    for v in range(V):
        fm.mask_device(shape, shape, A[v], 32, 1.0, mask.ptr)
        weights.append(mask.download())
    registered = [fm.DeviceArray(shape, np.uint16) for _ in range(V)]
    out = fm.DeviceArray(shape, np.uint16)
    log(f"[bench] data ready in {time.time() - t0:.1f}s on {name.value.decode()}")

    t0 = time.time()
    dec = fm.Deconvolver(H, W, weights, [k.copy() for k in kernels], psf_type=0, bits=16, batch_planes=Z)
    info = dec.info()
    log(f"[bench] plan {info} in {time.time() - t0:.1f}s")
    t0 = time.time()
    tuned = dec.autotune(planes=64)
    log(f"[bench] autotune in {time.time() - t0:.1f}s: " +
        ", ".join(f"{k}=v{c} {['%.3f' % x if x is not None else '-' for x in ts]}" for k, (c, ts) in tuned.items()))
    stream = fm.Stream()

    def step():
        for v in range(V):
            fm.transform_device(acquired[v].ptr, shape, registered[v].ptr, shape, A[v], stream=stream.handle)
        dec.run_device([r.ptr for r in registered], Z, ITERS, out_ptr=out.ptr, stream=stream.handle)

    for i in range(args.warmup):
        t = time.time()
        step()
        stream.synchronize()
        log(f"[bench] warmup {i}: {time.time() - t:.3f}s")

    fm.device_synchronize()
    with fm.KernelTiming() as kt:
        t = time.perf_counter()
        for i in range(args.steps):
            step()
        stream.synchronize()
        elapsed = time.perf_counter() - t
        timing = {k: kt.read(k) for k in ("column", "ratio", "update", "ingest", "egress", "transform")}
    ms_per_step = 1000.0 * elapsed / args.steps
    vvi = V * W * H * Z * ITERS
    value = vvi / (elapsed / args.steps)
    log(f"[bench] {args.steps} steps in {elapsed:.3f}s -> {ms_per_step:.1f} ms/step, {value:.4g} view-voxel it/s")

    fh, fw = info["fft_h"], info["fft_w"]
    S, P = fh * (fw // 2 + 1), fh * fw
    per_plane = {"column": 16 * S, "ratio": 16 * S + 2 * P, "update": 16 * S + 8 * P}
    # Each launch covers one cache-resident chunk of planes; the library times
    # the launches of every 8th chunk (event pairs around every launch would
    # add microseconds between kernels), all of them full chunks here.
    planes_per_launch = min(Z, info["batch"], info["chunk"])
    assert Z % planes_per_launch == 0, "timed chunks must be full"
    kstats = {}
    for k, b in per_plane.items():
        ms, n = timing[k]
        if n:
            avg = ms / n / 1000.0
            kstats[k] = dict(total_ms=ms, launches=n, avg_ms=1000 * avg, planes_per_launch=planes_per_launch,
                             achieved_gbs=b * planes_per_launch / avg / 1e9)
    dom = max(kstats, key=lambda k: kstats[k]["total_ms"])
    B_RL = 64 * S + 10 * P
    rl_ms = sum(kstats[k]["total_ms"] for k in kstats)
    # plane-view-iterations behind the timed launches: one ratio launch each
    iter_gbs = B_RL * kstats["ratio"]["launches"] * planes_per_launch / (rl_ms / 1000.0) / 1e9
    tr_ms, tr_n = timing["transform"]
    tr_gbs = (2 * Z * H * W * 2) * tr_n / (tr_ms / 1000.0) / 1e9 if tr_n else None
    for k, s in kstats.items():
        log(f"[bench] {k:7s} {s['launches']:5d} timed launches of {s['planes_per_launch']} planes, avg "
            f"{1000 * s['avg_ms']:.1f} us  {s['achieved_gbs']:.0f} GB/s")
    log(f"[bench] RL chain (B_RL = 64S+10P = {B_RL / 1e6:.2f} MB/view-iter-plane): {iter_gbs:.0f} GB/s; "
        f"transform {tr_gbs and round(tr_gbs)} GB/s")

    cpu = None
    if not args.no_cpu:
        import oracle
        zc = min(args.cpu_planes, Z)
        z0 = Z // 2 - zc // 2
        vols = [registered[v].download()[z0:z0 + zc] for v in range(V)]
        data = np.ascontiguousarray(np.stack(vols, axis=1))  # (zc, V, H, W)
        del vols
        kern = np.stack(kernels)
        wts = np.stack(weights)
        nt = cpu_threads()
        t = time.perf_counter()
        oracle.rl_planes(data, wts, kern.copy(), 0, 0, threads=nt)
        t_setup = time.perf_counter() - t
        t = time.perf_counter()
        oracle.rl_planes(data, wts, kern.copy(), 0, 1, threads=nt)
        t_one = time.perf_counter() - t
        t_iter = max(t_one - t_setup, 1e-9)
        cpu = {"value": V * W * H * zc * 1 / t_iter, "unit": "view-voxel RL iterations/s", "cores": nt,
               "kind": "port",
               "sample": f"FP64 oracle, {zc} planes x {V} views x 1 iteration of the same registered volume "
                         f"(FFT {fh}x{fw}), OpenMP over planes; plan setup ({t_setup:.1f}s) subtracted"}
        log(f"[bench] cpu baseline {cpu['value']:.4g} view-voxel it/s on {nt} threads ({t_iter:.1f}s)")

    for r in registered + acquired + [out]:
        r.free()
    dec.close()
    batched = None
    if not args.no_batch:
        batched = batched_workload(fm, synth)

    full = None
    if not args.no_full:
        full = full_size_leg(fm, synth)

    learned = None
    if not args.no_learned and not args.no_cpu:
        import oracle
        learned = learned_leg(fm, oracle, ntrain=args.train_samples, epochs=args.epochs)

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
            "workload": "affine transform + plain multi-view RL, 4 views uint16 1024x1024x512, "
                        "41x41 PSFs, 30 iterations, FFT 1536x1536",
            "views": V, "shape_zyx": [Z, H, W], "kernel": K, "iterations": ITERS, "fft": [fh, fw],
            "planes_per_launch": planes_per_launch, "device": name.value.decode(),
            "fft_plan_variants": {k: c for k, (c, _) in tuned.items()},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": dk["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dk["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": measured_traffic(fh, fw, dom, planes_per_launch),
        },
        "rl_iteration_roofline": {"achieved": iter_gbs, "frac": iter_gbs / HBM_PEAK_GBS, "unit": "GB/s",
                                  "bytes_per_view_iter_plane": B_RL},
        "kernels": kstats,
        "transform_gbs": tr_gbs,
        "cpu_baseline": cpu,
        "batched": batched,
        "full_size": full,
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
