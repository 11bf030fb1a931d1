"""CPU restatement of the learned predictor's network -- TEST INFRASTRUCTURE.

Only the tests, the smoke check and the benchmark's CPU-baseline leg use
this module, as the checker of the HIP kernels (no product path imports it).
The reference has no learned predictor (its deformation predictor is the
previous pass, gpu_process.py:620-625), so there is no reference algorithm
to follow: this file states the network the library builds -- architecture,
parameter layout, forward, weighted mean-square loss, backward, Adam and the
input preparation -- in plain NumPy (float64 unless asked to emulate BF16).

Pins (tests/test_cnn_oracle_pins.py): the convolution against a direct
seven-loop sum, the full backward against central finite differences, the
input preparation against explicit loops, and Adam's first step against its
closed form lr * g / (|g| + eps).
"""

from __future__ import annotations

import numpy as np

# (source 0, source 1 or -1, x2 upsampling of source 0, out channels, stride, ReLU);
# buffer 0 is the input (2 channels), buffer l+1 the output of layer l.
ARCH = (
    (0, -1, 0, 32, 1, 1), (1, -1, 0, 48, 2, 1), (2, -1, 0, 48, 1, 1), (3, -1, 0, 64, 2, 1),
    (4, -1, 0, 64, 1, 1), (5, -1, 0, 128, 2, 1), (6, -1, 0, 128, 1, 1), (7, 5, 1, 64, 1, 1),
    (8, -1, 0, 64, 1, 1), (9, 3, 1, 48, 1, 1), (10, -1, 0, 2, 1, 0),
)
FACTOR = 8


def layout():
    """Per layer ``(c0, c1, w_off, b_off)`` and the parameter count; weights
    are [co][3][3][c0+c1] then the bias [co], blocks padded to 4 floats."""
    ch = [2]
    off = 0
    out = []
    for s0, s1, up, co, stride, relu in ARCH:
        c0 = ch[s0]
        c1 = ch[s1] if s1 >= 0 else 0
        w_off = off
        off += co * 9 * (c0 + c1)
        b_off = off
        off += co
        off = (off + 3) // 4 * 4
        out.append((c0, c1, w_off, b_off))
        ch.append(co)
    return out, off


def init_params(seed=0):
    """He-normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    lay, n = layout()
    p = np.zeros(n, np.float32)
    for (c0, c1, w_off, b_off), spec in zip(lay, ARCH):
        co = spec[3]
        fan_in = 9 * (c0 + c1)
        p[w_off:w_off + co * fan_in] = rng.standard_normal(co * fan_in) * np.sqrt(2.0 / fan_in)
    return p


def weights(params, l):
    lay, _ = layout()
    c0, c1, w_off, b_off = lay[l]
    co = ARCH[l][3]
    w = params[w_off:w_off + co * 9 * (c0 + c1)].reshape(co, 3, 3, c0 + c1)
    return w, params[b_off:b_off + co]


def bf16(x):
    """Round to BF16 (nearest even) and back, in float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(x.shape)


def conv3x3(x, w, b, stride):
    """x (B, H, W, Ci), w (Co, 3, 3, Ci): zero padding 1, output H/stride."""
    bsz, h, wd, ci = x.shape
    ho, wo = h // stride, wd // stride
    xp = np.zeros((bsz, h + 2, wd + 2, ci), x.dtype)
    xp[:, 1:-1, 1:-1] = x
    out = np.zeros((bsz, ho, wo, w.shape[0]), np.result_type(x, w))
    for ky in range(3):
        for kx in range(3):
            tap = xp[:, ky:ky + stride * ho:stride, kx:kx + stride * wo:stride, :]
            out += tap @ w[:, ky, kx, :].T
    return out + b


def conv3x3_backward(x, w, stride, dy):
    """Gradients (dx, dw, db) of conv3x3 for output gradient dy."""
    bsz, h, wd, ci = x.shape
    ho, wo = dy.shape[1:3]
    xp = np.zeros((bsz, h + 2, wd + 2, ci), np.result_type(x, dy))
    xp[:, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros(w.shape, np.result_type(w, dy))
    flat = dy.reshape(-1, dy.shape[-1])
    for ky in range(3):
        for kx in range(3):
            sl = (slice(None), slice(ky, ky + stride * ho, stride), slice(kx, kx + stride * wo, stride))
            dw[:, ky, kx, :] = flat.T @ xp[sl].reshape(-1, ci)
            dxp[sl] += dy @ w[:, ky, kx, :]
    return dxp[:, 1:-1, 1:-1], dw, flat.sum(axis=0)


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def layer_input(bufs, l):
    s0, s1, up = ARCH[l][0], ARCH[l][1], ARCH[l][2]
    a = up2(bufs[s0]) if up else bufs[s0]
    return np.concatenate([a, bufs[s1]], axis=-1) if s1 >= 0 else a


def forward(params, x, emulate_bf16=False):
    """x (B, H, W, 2).  Returns every buffer: bufs[0] = x, bufs[l+1] = output
    of layer l (NHWC), the last one the head output.  With emulate_bf16 the
    weights and every stored activation are rounded to BF16 (the device's
    storage), products and sums stay in float32."""
    dt = np.float32 if emulate_bf16 else np.float64
    bufs = [np.asarray(x, dt)]
    for l, (s0, s1, up, co, stride, relu) in enumerate(ARCH):
        w, b = weights(np.asarray(params, dt), l)
        if emulate_bf16:
            w = bf16(w)
        y = conv3x3(layer_input(bufs, l), w, b, stride)
        if relu:
            y = np.maximum(y, 0)
        if emulate_bf16 and l < len(ARCH) - 1:
            y = bf16(y)
        bufs.append(y)
    return bufs


def planar(out):
    """NHWC (B, h, w, 2) -> per-image planar (B, 2, h, w)."""
    return np.ascontiguousarray(np.transpose(out, (0, 3, 1, 2)))


def loss_and_grad(out_planar, target, weight=None):
    """L = sum wt (o - y)^2 / count over (B, 2, h, w); returns (L, dL/do)."""
    d = out_planar - target
    wt = np.ones(out_planar.shape[:1] + out_planar.shape[2:]) if weight is None else weight
    wt = wt[:, None, :, :]
    count = out_planar.size
    return float(np.sum(wt * d * d) / count), 2.0 * wt * d / count


def backward(params, bufs, dout_planar):
    """Flat parameter gradient for the head-output gradient (planar)."""
    lay, n = layout()
    grads = np.zeros(n)
    g = [None] * len(bufs)
    g[-1] = np.transpose(dout_planar, (0, 2, 3, 1))
    for l in range(len(ARCH) - 1, -1, -1):
        s0, s1, up, co, stride, relu = ARCH[l]
        c0, c1, w_off, b_off = lay[l]
        dy = g[l + 1]
        if relu:
            dy = dy * (bufs[l + 1] > 0)
        w, _ = weights(np.asarray(params, np.float64), l)
        dx, dw, db = conv3x3_backward(layer_input(bufs, l), w, stride, dy)
        grads[w_off:w_off + dw.size] = dw.reshape(-1)
        grads[b_off:b_off + co] = db
        if s0 == 0:
            continue
        d0 = dx[..., :c0]
        if up:
            b_, h_, w_, c_ = d0.shape
            d0 = d0.reshape(b_, h_ // 2, 2, w_ // 2, 2, c_).sum(axis=(2, 4))
        g[s0] = d0 if g[s0] is None else g[s0] + d0
        if s1 >= 0:
            d1 = dx[..., c0:]
            g[s1] = d1 if g[s1] is None else g[s1] + d1
    return grads


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def prepare_input(frame_a, frame_b, f=FACTOR):
    """f x f box means of each frame, standardised to zero mean and unit
    variance, stacked as (H/f, W/f, 2)."""
    out = []
    for fr in (frame_a, frame_b):
        h, w = fr.shape
        box = np.asarray(fr, np.float64)[:h // f * f, :w // f * f].reshape(h // f, f, w // f, f).mean(axis=(1, 3))
        sd = box.std()
        out.append((box - box.mean()) / (sd if sd > 0 else 1.0))
    return np.stack(out, axis=-1)
