"""The warm-start network's GPU kernels against the FP64 oracle
(oracle/net.py, run with the same bfloat16 roundings as the GPU's storage):
the MFMA operand layout, the forward pass, every parameter gradient, one
Adam step, and tile-wise inference against whole-volume inference."""
import numpy as np
import pytest

from conftest import relL2
from oracle import net as onet

pytestmark = pytest.mark.gpu


def flat(params):
    return np.concatenate([np.concatenate([w.reshape(-1), b.reshape(-1)]) for w, b in params]).astype(np.float32)


def unflat(v):
    out, o = [], 0
    for ws, bs in onet.param_shapes():
        nw, nb = int(np.prod(ws)), int(np.prod(bs))
        out.append([v[o:o + nw].reshape(ws).astype(np.float64), v[o + nw:o + nw + nb].reshape(bs).astype(np.float64)])
        o += nw + nb
    return out


def test_mfma_operand_layout(fm):
    rng = np.random.default_rng(0)
    A = onet.bf16(rng.normal(size=(32, 16)))
    B = onet.bf16(rng.normal(size=(16, 32)))
    net = fm.Net(0)
    dA = fm.DeviceArray.from_host(fm.to_bf16_bits(A))
    dB = fm.DeviceArray.from_host(fm.to_bf16_bits(B))
    dD = fm.DeviceArray((32, 32), np.float32)
    assert net.L.fmvd_net_mfma_layout_test(dA.ptr, dB.ptr, dD.ptr) == 0
    fm.check()
    np.testing.assert_allclose(dD.download(), A @ B, rtol=1e-5, atol=1e-5)


def make_problem(seed, shape=(2, 5, 9, 40)):
    rng = np.random.default_rng(seed)
    params = onet.init_params(seed)
    for p in params:
        p[0] *= 2.0
        p[1][...] = rng.normal(0.0, 0.05, p[1].shape)
    params = unflat(flat(params))  # FP32-representable master weights
    N, D, H, W = shape
    x = onet.bf16(rng.uniform(0.2, 1.5, size=(N, D, H, W, 2)))
    t = rng.uniform(0.2, 1.5, size=(N, D, H, W)).astype(np.float32).astype(np.float64)
    return params, x, t


def oracle_batch(params, x, t):
    """Oracle loss/grads for a batch: the mean over all samples' voxels."""
    N = x.shape[0]
    loss, grads, ys = 0.0, None, []
    for n in range(N):
        l, g, y = onet.loss_and_grads(params, x[n], t[n], quant=True)
        loss += l / N
        ys.append(y)
        g = [[gw / N, gb / N] for gw, gb in g]
        grads = g if grads is None else [[a[0] + b[0], a[1] + b[1]] for a, b in zip(grads, g)]
    return loss, grads, np.stack(ys)


def run_gpu(fm, params, x, t):
    net = fm.Net(1)
    net.set_params(flat(params))
    dx = fm.DeviceArray.from_host(fm.to_bf16_bits(x))
    dt = fm.DeviceArray.from_host(t.astype(np.float32))
    dy = fm.DeviceArray(t.shape, np.float32)
    net.forward(dx, t.shape, dy)
    fm.device_synchronize()
    y = dy.download()
    loss = net.backward(dx, dt, t.shape)
    return net, y, loss, net.grads()


def test_forward_and_gradients_match_oracle(fm):
    params, x, t = make_problem(7)
    net, y, loss, g = run_gpu(fm, params, x, t)
    loss_o, grads_o, y_o = oracle_batch(params, x, t)
    assert relL2(y, y_o) < 2e-3, relL2(y, y_o)
    assert abs(loss - loss_o) < 2e-3 * loss_o
    go = unflat(flat(grads_o).astype(np.float64)) if False else grads_o
    gg = unflat(g)
    for l in range(onet.LAYERS):
        for i in range(2):
            r = relL2(gg[l][i], go[l][i])
            print(f"layer {l + 1} {'W' if i == 0 else 'b'}: relL2 {r:.2e}")
            assert r < 2e-2, (l, i, r)


def test_adam_step_matches_oracle(fm):
    params, x, t = make_problem(11, shape=(1, 4, 6, 33))
    net, _, _, g = run_gpu(fm, params, x, t)
    net.adam(1e-3)
    got = net.params()
    p = unflat(flat(params))
    m = [[np.zeros_like(a), np.zeros_like(b)] for a, b in p]
    v = [[np.zeros_like(a), np.zeros_like(b)] for a, b in p]
    onet.adam_step(p, unflat(g), m, v, 1, lr=1e-3)
    np.testing.assert_allclose(got, flat(p), rtol=0, atol=2e-6)


def test_training_reduces_loss(fm):
    """80 Adam steps at lr 2e-4 on a smoothing task; the FP64 oracle trained
    the same way falls to ~7% of its initial loss."""
    rng = np.random.default_rng(5)
    N, D, H, W = 2, 6, 16, 32
    x = onet.bf16(rng.uniform(0.2, 1.5, size=(N, D, H, W, 2)))
    m = 0.5 * (x[..., 0] + x[..., 1])
    mp = np.pad(m, ((0, 0), (0, 0), (0, 0), (1, 1)))
    t = (0.5 * m + 0.25 * (mp[..., 2:] + mp[..., :-2])).astype(np.float32)
    net = fm.Net(3)
    dx = fm.DeviceArray.from_host(fm.to_bf16_bits(x))
    dt = fm.DeviceArray.from_host(t)
    losses = []
    for _ in range(80):
        losses.append(net.backward(dx, dt, t.shape))
        net.adam(2e-4)
    print("losses", [round(l, 5) for l in losses[::10]], losses[-1])
    assert losses[-1] < 0.5 * losses[0]
    assert np.isfinite(losses).all()


def test_tiled_inference_equals_whole_volume(fm):
    rng = np.random.default_rng(9)
    D, H, W, V = 20, 70, 90, 3
    views = [fm.DeviceArray.from_host(rng.integers(50, 900, size=(D, H, W)).astype(np.uint16)) for _ in range(V)]
    net = fm.Net(4)
    whole = fm.DeviceArray((D, H, W), np.float32)
    tiled = fm.DeviceArray((D, H, W), np.float32)
    net.infer(views, 300.0, whole, tile=(D, H, W, 16))
    net.infer(views, 300.0, tiled, tile=(8, 24, 32, 8))
    a, b = whole.download(), tiled.download()
    assert np.isfinite(a).all() and a.min() >= 1e-3
    assert np.array_equal(a, b)
