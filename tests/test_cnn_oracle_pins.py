"""Pins of the network's CPU restatement (oracle/cnn_oracle.py): a direct
seven-loop convolution, central finite differences of the whole loss, the
input preparation by explicit loops, Adam's closed-form first step."""

import numpy as np
import pytest

from oracle import cnn_oracle as C


def _direct_conv(x, w, b, stride):
    bsz, h, wd, ci = x.shape
    co = w.shape[0]
    ho, wo = h // stride, wd // stride
    out = np.zeros((bsz, ho, wo, co))
    for n in range(bsz):
        for y in range(ho):
            for xx in range(wo):
                for o in range(co):
                    s = b[o]
                    for ky in range(3):
                        for kx in range(3):
                            yi, xi = stride * y + ky - 1, stride * xx + kx - 1
                            if 0 <= yi < h and 0 <= xi < wd:
                                s += np.dot(w[o, ky, kx, :], x[n, yi, xi, :])
                    out[n, y, xx, o] = s
    return out


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_matches_direct_sum(stride):
    rng = np.random.default_rng(stride)
    x = rng.standard_normal((2, 6, 8, 3))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    assert np.allclose(C.conv3x3(x, w, b, stride), _direct_conv(x, w, b, stride), atol=1e-12)


def test_conv_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 6, 6, 2))
    w = rng.standard_normal((3, 3, 3, 2))
    b = rng.standard_normal(3)
    dy = rng.standard_normal((1, 3, 3, 3))
    dx, dw, db = C.conv3x3_backward(x, w, 2, dy)
    f = lambda x_, w_, b_: float(np.sum(C.conv3x3(x_, w_, b_, 2) * dy))
    eps = 1e-6
    for idx in [(0, 0, 0, 0), (0, 3, 5, 1), (0, 5, 5, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += eps
        xm[idx] -= eps
        assert abs((f(xp, w, b) - f(xm, w, b)) / (2 * eps) - dx[idx]) < 1e-6
    for idx in [(0, 0, 0, 0), (2, 1, 2, 1)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps
        wm[idx] -= eps
        assert abs((f(x, wp, b) - f(x, wm, b)) / (2 * eps) - dw[idx]) < 1e-6
    assert np.allclose(db, dy.sum(axis=(0, 1, 2)))


def test_layout_parameter_count():
    lay, n = C.layout()
    assert len(lay) == len(C.ARCH)
    raw = sum(spec[3] * 9 * (c0 + c1) + spec[3] for (c0, c1, _, _), spec in zip(lay, C.ARCH))
    assert raw == 518226 and n >= raw
    # decoder layers see the upsampled tensor plus the skip
    assert lay[7][:2] == (128, 64) and lay[9][:2] == (64, 48)


def test_network_gradient_matches_finite_differences():
    """Whole forward + loss + backward at a 16 x 16 input (output 8 x 8)."""
    rng = np.random.default_rng(4)
    p = C.init_params(5).astype(np.float64)
    lay, n = C.layout()
    for (_, _, _, b_off), spec in zip(lay, C.ARCH):
        p[b_off:b_off + spec[3]] = rng.uniform(0.05, 0.2, spec[3])  # keep ReLUs away from their kink
    x = rng.standard_normal((2, 16, 16, 2))
    target = rng.standard_normal((2, 2, 8, 8))
    weight = rng.uniform(0.5, 1.5, (2, 8, 8))

    def loss(pp):
        bufs = C.forward(pp, x)
        return C.loss_and_grad(C.planar(bufs[-1]), target, weight)[0]

    bufs = C.forward(p, x)
    L, dout = C.loss_and_grad(C.planar(bufs[-1]), target, weight)
    g = C.backward(p, bufs, dout)
    eps = 1e-6
    checked = 0
    for l, ((c0, c1, w_off, b_off), spec) in enumerate(zip(lay, C.ARCH)):
        for idx in list(rng.choice(spec[3] * 9 * (c0 + c1), 3, replace=False) + w_off) + [b_off]:
            pp, pm = p.copy(), p.copy()
            pp[idx] += eps
            pm[idx] -= eps
            fd = (loss(pp) - loss(pm)) / (2 * eps)
            assert abs(fd - g[idx]) <= 1e-5 * max(1.0, abs(fd)) + 1e-8, (l, idx, fd, g[idx])
            checked += 1
    assert checked == 4 * len(C.ARCH)


def test_prepare_input_matches_loops():
    rng = np.random.default_rng(6)
    a = rng.integers(0, 4096, (32, 48)).astype(np.uint16)
    b = rng.integers(0, 4096, (32, 48)).astype(np.uint16)
    got = C.prepare_input(a, b)
    for k, fr in enumerate((a, b)):
        box = np.zeros((4, 6))
        for i in range(4):
            for j in range(6):
                box[i, j] = sum(float(fr[8 * i + y, 8 * j + x]) for y in range(8) for x in range(8)) / 64
        ref = (box - box.mean()) / box.std()
        assert np.allclose(got[..., k], ref, atol=1e-12)


def test_adam_first_step_closed_form():
    g = np.array([0.5, -2.0, 1e-3])
    p, m, v = C.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    assert np.allclose(p, -0.01 * g / (np.abs(g) + 1e-8))


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -3.14159], np.float32)
    y = C.bf16(x)
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.0 + 2 ** -6 and abs(y[3] + 3.140625) < 1e-7
