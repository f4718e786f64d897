"""Batched RL over many volumes of uneven size (fmvd_batch.h): every
volume's result must be bit-identical to a single-volume plan run on the same
device data with the same FFT plan variants, graph replays must equal the
eagerly enqueued iterations, and one volume is checked against the FP64
oracle directly."""
import numpy as np
import pytest

from gpu_util import check_rl, make_case

pytestmark = pytest.mark.gpu

# Z, V, dh, dw, k, psf  -> FFT class (fftH x fftW)
VOLUMES = [
    (3, 2, 150, 170, 31, 0),   # 256 x 256
    (2, 3, 200, 180, 31, 1),   # 256 x 256
    (4, 2, 230, 200, 31, 2),   # 512 x 256
    (2, 4, 130, 330, 41, 0),   # 256 x 512
    (3, 2, 250, 240, 31, 3),   # 512 x 512
    (1, 3, 400, 380, 41, 0),   # 512 x 512
    (2, 2, 500, 520, 41, 0),   # 1024 x 1024
    (2, 2, 151, 171, 31, 0),   # 256 x 256, odd data width
]


def make_volumes(fm, vols, bits=16, seed0=100):
    out = []
    dt = np.uint16 if bits == 16 else np.uint8
    for i, (Z, V, dh, dw, k, psf) in enumerate(vols):
        data, kernels, weights = make_case(Z, V, dh, dw, k, k, bits, seed=seed0 + i, beads=30)
        views = [fm.DeviceArray.from_host(np.ascontiguousarray(data[:, v].astype(dt))) for v in range(V)]
        out.append(dict(data=data, raw_kernels=kernels, weights=list(weights), views=views, psf_type=psf,
                        kernels=[kk.copy() for kk in kernels],
                        out=fm.DeviceArray((Z, dh, dw), dt), est=fm.DeviceArray((Z, dh, dw), np.float32)))
    return out


def single(fm, v, iters, bits=16, variants=None):
    Z, dh, dw = v["views"][0].shape
    dt = np.uint16 if bits == 16 else np.uint8
    o = fm.DeviceArray((Z, dh, dw), dt)
    e = fm.DeviceArray((Z, dh, dw), np.float32)
    with fm.Deconvolver(dh, dw, v["weights"], [k.copy() for k in v["raw_kernels"]], psf_type=v["psf_type"],
                        bits=bits) as dec:
        if variants is not None:
            dec.set_variants(variants)
        dec.run_device([d.ptr for d in v["views"]], Z, iters, out_ptr=o.ptr, est_ptr=e.ptr)
        fm.device_synchronize()
    return o.download(), e.download()


@pytest.mark.parametrize("bits", [16, 8])
def test_batch_matches_single_volume_plans(fm, bits):
    vols = make_volumes(fm, VOLUMES, bits)
    iters = 4
    print(f"\n[batch] {len(vols)} volumes bits={bits}", flush=True)
    with fm.Batch(vols, bits=bits) as b:
        classes = b.info()
        assert sorted((c["fft_h"], c["fft_w"]) for c in classes) == [(256, 256), (256, 512), (512, 256),
                                                                       (512, 512), (1024, 1024)]
        b.run(iters, use_graph=True)
        fm.device_synchronize()
        for i, v in enumerate(vols):
            o, e = single(fm, v, iters, bits)
            assert np.array_equal(v["out"].download(), o), i
            assert np.array_equal(v["est"].download(), e), i


def test_chunked_batch_matches_single_volume_plans(fm, monkeypatch):
    """A 2 MiB cache budget splits volumes into plane ranges spread over many
    chunks (one plane per chunk at 1024 x 1024); results must not change."""
    monkeypatch.setenv("FMVD_CHUNK_BYTES", str(2 << 20))
    vols = make_volumes(fm, VOLUMES, seed0=300)
    iters = 3
    print(f"\n[batch chunked] {len(vols)} volumes", flush=True)
    with fm.Batch(vols) as b:
        b.run(iters, use_graph=True)
        fm.device_synchronize()
        graph = [(v["out"].download(), v["est"].download()) for v in vols]
        b.run(iters, use_graph=False)
        fm.device_synchronize()
        for i, v in enumerate(vols):
            o, e = single(fm, v, iters)
            assert np.array_equal(graph[i][0], o) and np.array_equal(graph[i][1], e), i
            assert np.array_equal(v["out"].download(), o) and np.array_equal(v["est"].download(), e), i


def test_graph_replays_equal_eager(fm):
    vols = make_volumes(fm, VOLUMES[:4], seed0=200)
    with fm.Batch(vols) as b:
        b.run(3, use_graph=False)
        fm.device_synchronize()
        ref = [(v["out"].download(), v["est"].download()) for v in vols]
        scratch = fm.DeviceArray((1 << 20,), np.uint8)
        for rep in range(3):
            for v in vols:  # clear the outputs between replays with eager work
                v["out"].fill_zero()
                v["est"].fill_zero()
            scratch.fill_zero()
            fm.device_synchronize()
            b.run(3, use_graph=True)
            fm.device_synchronize()
            for i, v in enumerate(vols):
                assert np.array_equal(v["out"].download(), ref[i][0]), (rep, i)
                assert np.array_equal(v["est"].download(), ref[i][1]), (rep, i)


def test_autotuned_batch_matches_oracle_and_single_plan(fm, orc):
    vols = make_volumes(fm, [(2, 2, 150, 170, 31, 0), (2, 3, 130, 330, 41, 2)], seed0=300)
    iters = 3
    with fm.Batch(vols) as b:
        chosen = b.autotune()
        classes = b.info()
        print(f"\n[batch autotune] {classes} -> {chosen}", flush=True)
        b.run(iters, use_graph=True)
        fm.device_synchronize()
        for i, v in enumerate(vols):
            c = [k for k, cl in enumerate(classes)
                 if (cl["fft_h"], cl["fft_w"]) == ({0: (256, 256), 1: (256, 512)}[i])][0]
            col, ratio, upd = chosen[c]
            o, e = single(fm, v, iters, variants=[col, ratio, upd, 0, 0])
            assert np.array_equal(v["out"].download(), o), i
            assert np.array_equal(v["est"].download(), e), i
        v = vols[0]
        out_o, est_o = orc.rl_planes(v["data"], np.stack(v["weights"]), v["raw_kernels"].copy(), v["psf_type"],
                                     iters, threads=8)
        check_rl(v["out"].download(), v["est"].download(), out_o, est_o)
