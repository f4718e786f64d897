"""FP64 numpy restatement of the warm-start network (forward, backward, Adam).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's checking legs,
never by the product.  The network has no counterpart in the reference (it is
an addition of this build); this module fixes its definition, and the GPU
kernels (gpu_deconvolution_amd/csrc/net_*.hip) are checked against it.
Pinned by tests/test_oracle_net.py: the convolution against
scipy.ndimage.correlate, every gradient against central finite differences,
Adam against its closed-form first step.

Network (6 layers, 32 channels, 3x3x3, zero padding):
  x   : (D, H, W, 2)  channels = mean of the even-index views, mean of the
                      odd-index views, both divided by the sample scale s
  h1  = relu(conv(x;  W1 (27, 32, 2))  + b1)
  h_l = h_{l-1} + relu(conv(h_{l-1}; W_l (27, 32, 32)) + b_l),  l = 2..5
  y   = conv(h5; W6 (27, 1, 32)) + b6 + (x0 + x1) / 2
  loss = mean((y - t)^2),  t = target / s
Tap index k = (dz * 3 + dy) * 3 + dx reads input voxel (z+dz-1, y+dy-1, x+dx-1).
With quant=True the weights, x and the stored activations (h1, r_l = relu(a_l),
h_l, and the masked gradients g_l) are rounded to bfloat16 (round to nearest
even) where the GPU stores them in bfloat16; all arithmetic stays FP64.
"""
from __future__ import annotations

import numpy as np

C = 32
LAYERS = 6


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def param_shapes():
    s = [((27, C, 2), (C,))]
    s += [((27, C, C), (C,))] * 4
    s += [((27, 1, C), (1,))]
    return s


def init_params(seed: int = 0):
    """He-style init scaled down for the residual blocks; last layer small."""
    rng = np.random.default_rng(seed)
    ps = []
    for l, (ws, bs) in enumerate(param_shapes()):
        fan_in = 27 * ws[2]
        std = np.sqrt(2.0 / fan_in) * (0.5 if 1 <= l <= 4 else 1.0) * (0.1 if l == 5 else 1.0)
        ps.append([rng.normal(0.0, std, ws), np.zeros(bs)])
    return ps


def _pad(x):
    return np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))


def conv(x: np.ndarray, w: np.ndarray) -> np.ndarray:
    """out[v, co] = sum_{k, ci} x[v + off(k), ci] w[k, co, ci], zero padding."""
    D, H, W, _ = x.shape
    xp = _pad(x)
    out = np.zeros((D, H, W, w.shape[1]))
    for k in range(27):
        dz, dy, dx = k // 9, (k // 3) % 3, k % 3
        out += xp[dz:dz + D, dy:dy + H, dx:dx + W, :] @ w[k].T
    return out


def conv_t(g: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Adjoint of conv in its input: d[v, ci] = sum_{k, co} g[v - off(k), co] w[k, co, ci]."""
    D, H, W, _ = g.shape
    out = np.zeros((D + 2, H + 2, W + 2, w.shape[2]))
    for k in range(27):
        dz, dy, dx = k // 9, (k // 3) % 3, k % 3
        out[dz:dz + D, dy:dy + H, dx:dx + W, :] += g @ w[k]
    return out[1:-1, 1:-1, 1:-1, :]


def conv_wgrad(x: np.ndarray, g: np.ndarray) -> np.ndarray:
    """dW[k, co, ci] = sum_v g[v, co] x[v + off(k), ci]."""
    D, H, W, _ = x.shape
    xp = _pad(x)
    out = np.zeros((27, g.shape[3], x.shape[3]))
    gf = g.reshape(-1, g.shape[3])
    for k in range(27):
        dz, dy, dx = k // 9, (k // 3) % 3, k % 3
        out[k] = gf.T @ xp[dz:dz + D, dy:dy + H, dx:dx + W, :].reshape(-1, x.shape[3])
    return out


def forward(params, x: np.ndarray, quant: bool = False):
    """Returns (y, cache).  x: (D, H, W, 2)."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    W = [(q(w), np.asarray(b, np.float64)) for w, b in params]
    x = q(x)
    a1 = conv(x, W[0][0]) + W[0][1]
    h = q(np.maximum(a1, 0.0))
    cache = {"x": x, "h": [h], "r": [h], "W": W}
    for l in range(1, 5):
        a = conv(h, W[l][0]) + W[l][1]
        r = q(np.maximum(a, 0.0))
        h = q(h + r)
        cache["r"].append(r)
        cache["h"].append(h)
    y = conv(h, W[5][0])[..., 0] + W[5][1][0] + 0.5 * (x[..., 0] + x[..., 1])
    return y, cache


def loss_and_grads(params, x, t, quant: bool = False):
    """MSE loss and its gradients for every (W, b)."""
    q = bf16 if quant else (lambda a: a)
    y, c = forward(params, x, quant)
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    dy = (2.0 / n) * (y - t)
    grads = [None] * LAYERS
    W = c["W"]
    h = c["h"]
    grads[5] = [conv_wgrad(h[4], dy[..., None]), np.array([dy.sum()])]
    dh = conv_t(dy[..., None], W[5][0])
    for l in range(4, 0, -1):
        g = q(dh * (c["r"][l] > 0))
        grads[l] = [conv_wgrad(h[l - 1], g), g.reshape(-1, C).sum(0)]
        dh = dh + conv_t(g, W[l][0])
    g = q(dh * (c["r"][0] > 0))
    grads[0] = [conv_wgrad(c["x"], g), g.reshape(-1, C).sum(0)]
    return loss, grads, y


def adam_step(params, grads, m, v, step: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """In-place Adam (Kingma & Ba 2015, bias-corrected); step starts at 1."""
    for p, g, mm, vv in zip(params, grads, m, v):
        for i in range(2):
            mm[i][...] = b1 * mm[i] + (1 - b1) * g[i]
            vv[i][...] = b2 * vv[i] + (1 - b2) * g[i] ** 2
            mh = mm[i] / (1 - b1 ** step)
            vh = vv[i] / (1 - b2 ** step)
            p[i][...] = p[i] - lr * mh / (np.sqrt(vh) + eps)


def views_to_input(views: np.ndarray, scale: float) -> np.ndarray:
    """views: (V, D, H, W) samples -> (D, H, W, 2) network input."""
    v = np.asarray(views, np.float64)
    even = v[0::2].mean(0)
    odd = v[1::2].mean(0) if v.shape[0] > 1 else even
    return np.stack([even, odd], -1) / scale
