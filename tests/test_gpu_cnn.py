"""The learned predictor's HIP kernels against the CPU oracle
(oracle/cnn_oracle.py): layout, forward activations of every layer, the
loss, every parameter gradient, the Adam update, the input preparation, a
short training run, and the learned path of a plan (predictor plumbing and
the fallback rule).

Tolerances: activations and weights are BF16 on the device (8-bit mantissa,
relative rounding 2^-9), so the forward is compared with the oracle run on
the same BF16-rounded weights and stored activations (max error 2 % of each
buffer's range); gradients additionally round the output gradient to BF16 as
MFMA operands (relative L2 error per layer <= 3 %, cosine >= 0.999)."""

import numpy as np
import pytest

from oracle import cnn_oracle as C
from oracle import piv_oracle as O

pytestmark = pytest.mark.gpu


def _bits(x):
    return (C.bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def _net(frame=(512, 512), batch=2, params=None):
    from openpiv_gpu.learned import CoarseFlowNet
    return CoarseFlowNet(frame, max_batch=batch, params=params)


def _params(seed):
    rng = np.random.default_rng(seed)
    p = C.init_params(seed)
    lay, _ = C.layout()
    for (_, _, _, b_off), spec in zip(lay, C.ARCH):
        p[b_off:b_off + spec[3]] = rng.uniform(-0.05, 0.1, spec[3])
    return p


def _input(seed, batch=2, shape=(64, 64)):
    rng = np.random.default_rng(seed)
    return C.bf16(rng.standard_normal((batch,) + shape + (2,)).astype(np.float32))


def test_layout_matches_oracle():
    net = _net(batch=1)
    lay, n = C.layout()
    assert net.n_params == n
    for l, ((c0, c1, w_off, b_off), spec) in enumerate(zip(lay, C.ARCH)):
        s0, s1, up, co, stride, relu = spec
        assert tuple(net.layer_desc[l]) == (s0, s1, up, c0, c1, co, stride, relu)
        assert tuple(net.layer_offsets[l]) == (w_off, b_off)


def test_forward_matches_oracle():
    from openpiv_gpu import _native
    p = _params(1)
    net = _net(params=p)
    x = _input(2)
    dx = _native.DeviceBuffer.from_array(_bits(x))
    net.forward(dx.value, 2)
    ref = C.forward(p, x, emulate_bf16=True)
    for b in range(1, len(C.ARCH) + 1):
        got = net.get_buffer(b, 2)
        r = C.planar(ref[b]) if b == len(C.ARCH) else ref[b]
        scale = max(np.abs(r).max(), 1e-6)
        assert np.abs(got - r).max() <= 0.02 * scale, (b, np.abs(got - r).max(), scale)


def test_gradients_loss_and_adam_match_oracle():
    from openpiv_gpu import _native
    p = _params(3)
    net = _net(params=p)
    x = _input(4)
    rng = np.random.default_rng(5)
    target = rng.standard_normal((2, 2, 32, 32)).astype(np.float32)
    weight = rng.uniform(0.0, 1.0, (2, 32, 32)).astype(np.float32)
    dx = _native.DeviceBuffer.from_array(_bits(x))
    dt = _native.DeviceBuffer.from_array(target)
    dw = _native.DeviceBuffer.from_array(weight)
    loss = _native.DeviceBuffer((1,), np.float32)
    net.gradients(dx.value, dt.value, dw.value, 2, loss.value)
    g = net.get_grads()
    bufs = [np.asarray(b, np.float64) for b in C.forward(p, x, emulate_bf16=True)]
    L, dout = C.loss_and_grad(C.planar(bufs[-1]), target, weight)
    assert abs(float(loss.get()[0]) - L) <= 1e-3 * L
    ref = C.backward(p, bufs, dout)
    lay, _ = C.layout()
    for l, ((c0, c1, w_off, b_off), spec) in enumerate(zip(lay, C.ARCH)):
        co = spec[3]
        for lo, hi in ((w_off, w_off + co * 9 * (c0 + c1)), (b_off, b_off + co)):
            a, r = g[lo:hi].astype(np.float64), ref[lo:hi]
            rel = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)
            cos = float(a @ r / max(np.linalg.norm(a) * np.linalg.norm(r), 1e-30))
            assert rel <= 0.03 and cos >= 0.999, (l, lo, rel, cos)
    # one Adam step from zero moments
    lr = 1e-3
    net.adam_step(lr)
    p1 = net.get_params()
    ref_p, _, _ = C.adam_step(p.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, lr)
    assert np.abs(p1 - ref_p).max() <= 1e-6 + 1e-5 * lr


def test_training_reduces_loss():
    from openpiv_gpu import _native
    net = _net(params=C.init_params(7))
    x = _input(8)
    yy, xx = np.meshgrid(np.arange(32), np.arange(32), indexing="ij")
    target = np.stack([np.stack([np.sin(xx / 5.0 + k), np.cos(yy / 7.0 - k)]) for k in range(2)]).astype(np.float32)
    dx = _native.DeviceBuffer.from_array(_bits(x))
    dt = _native.DeviceBuffer.from_array(target)
    loss = _native.DeviceBuffer((1,), np.float32)
    losses = []
    for it in range(60):
        net.gradients(dx.value, dt.value, None, 2, loss.value)
        net.adam_step(1e-3)
        if it in (0, 59):
            losses.append(float(loss.get()[0]))
    assert losses[1] < 0.5 * losses[0], losses


def test_prepare_input_matches_oracle():
    from openpiv_gpu import _native
    rng = np.random.default_rng(9)
    fa = rng.integers(0, 4096, (512, 512)).astype(np.uint16)
    fb = (fa.astype(np.int32) // 2 + rng.integers(0, 100, fa.shape)).astype(np.uint16)
    net = _net(batch=1)
    out = _native.DeviceBuffer((64, 64, 2), np.uint16)
    net.prepare_input(_native.DeviceBuffer.from_array(fa), _native.DeviceBuffer.from_array(fb), out.value)
    got = (out.get().astype(np.uint32) << 16).view(np.float32)
    ref = C.prepare_input(fa, fb)
    assert np.abs(got - ref).max() <= 2 ** -8 * np.abs(ref).max() + 1e-6


def _constant_predictor_net(u, v):
    """Head weights zero, head bias (u, v): the network predicts a uniform shift."""
    p = C.init_params(11)
    lay, _ = C.layout()
    c0, c1, w_off, b_off = lay[-1]
    p[w_off:b_off] = 0.0
    p[b_off:b_off + 2] = (u, v)
    return _net(batch=1, params=p)


def test_learned_path_matches_oracle_last_pass():
    from openpiv_gpu import synthetic as S
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (512, 512)
    fa, fb = S.make_pair(shape, S.uniform_displacement(2.35, -1.6), S.ParticleModel(), 21)
    net = _constant_predictor_net(2.3, -1.5)
    piv = PIVGPU(shape, 16, search_size_iters=(1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float32",
                 predictor=net, max_reject_fraction=1.0)
    u, v = piv.run_host(fa, fb)
    fell_back, rejected = piv.last_fallback
    assert not fell_back
    # oracle: the last pass alone, predictor = the constant field on the network's nodes
    g_last = O.field_geometry(shape, 16, 0.5)
    nodes = np.arange(32) * 16 + 8.0
    xg, yg = np.meshgrid(nodes, nodes)
    # (nominal window 32 so that the oracle interpolates between the grids, gpu_process.py:732)
    g_net = O.Geometry(shape, 32, 16, 16, (32, 32), (-8, -8), xg.astype(np.float32), yg.astype(np.float32))
    s = O.OracleSettings(min_search_size=16, search_size_iters=(1, 1), overlap_ratio=0.5, median_tol=2.0)
    st = O.run_pass(fa, fb, s, g_last, g_net, np.full((32, 32), 2.3), np.full((32, 32), -1.5), True)
    assert np.abs(u - st.u_out).max() <= 1e-3 and np.abs(v - st.v_out).max() <= 1e-3
    assert rejected == int(st.val.sum())
    assert piv.outlier_counts(1) == (int(st.val.sum()), int(st.outliers.sum()))
    # second call replays the captured graph: identical output, and the
    # counters start from zero again
    u2, v2 = piv.run_host(fa, fb)
    assert np.array_equal(u, u2) and np.array_equal(v, v2)
    assert piv.last_fallback == (False, rejected)
    assert piv.outlier_counts(1) == (int(st.val.sum()), int(st.outliers.sum()))


def test_learned_path_falls_back_to_full_schedule():
    from openpiv_gpu import synthetic as S
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (512, 512)
    fa, fb = S.make_pair(shape, S.uniform_displacement(2.35, -1.6), S.ParticleModel(), 22)
    kw = dict(search_size_iters=(1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float32")
    plain = PIVGPU(shape, 16, **kw)
    up, vp = plain.run_host(fa, fb)
    # a predictor 25 px off: the last pass alone cannot recover, most vectors are rejected
    learned = PIVGPU(shape, 16, predictor=_constant_predictor_net(25.0, 25.0), max_reject_fraction=0.05, **kw)
    counts = [plain.outlier_counts(k) for k in range(2)]
    for _ in range(2):  # eager, then graph
        u, v = learned.run_host(fa, fb)
        fell_back, rejected = learned.last_fallback
        assert fell_back and rejected > 0.05 * u.size
        assert np.array_equal(u, up) and np.array_equal(v, vp)
        # the full-schedule run's counters are its own (cleared after the
        # learned attempt), equal to the plain plan's
        assert [learned.outlier_counts(k) for k in range(2)] == counts


def test_learned_runs_queued_async_equal_synchronous_runs():
    """piv_run_async on learned plans: pairs queued back to back, the
    fallback decided at piv_plan_synchronize.  Every queued pair equals its
    synchronous run bit for bit -- the learned result when the network's
    field passes the 5 % test, the full schedule's when it does not -- and
    the learned-run counters count both."""
    from openpiv_gpu import _native
    from openpiv_gpu import synthetic as S
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (512, 512)
    kw = dict(search_size_iters=(1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float32")
    pairs = [S.make_pair(shape, S.uniform_displacement(2.35, -1.6), S.ParticleModel(), 30 + i) for i in range(4)]
    frames = _native.DeviceBuffer.from_array(np.stack([np.stack(p) for p in pairs]).astype(np.uint8))
    for net, fell in ((_constant_predictor_net(2.3, -1.5), False), (_constant_predictor_net(25.0, 25.0), True)):
        sync = PIVGPU(shape, 16, predictor=net, max_reject_fraction=0.05, **kw)
        asy = PIVGPU(shape, 16, predictor=net, max_reject_fraction=0.05, **kw)
        ps, pa = sync._get_plan(_native.PIV_U8), asy._get_plan(_native.PIV_U8)
        m, n = sync.piv_fields[-1].field_shape
        outs = _native.DeviceBuffer((2, len(pairs), 2, m, n), np.float32)
        flags = _native.DeviceBuffer((2, len(pairs), 2, m, n), np.uint8)
        for rep in range(2):  # eager, then graph replays
            for i in range(len(pairs)):
                f, o, fl = frames.view(i), outs.view(0).view(i), flags.view(0).view(i)
                ps.run(f.view(0).value, f.view(1).value, True, o.view(0).value, o.view(1).value, fl.view(0).value,
                       fl.view(1).value, True)
                assert ps.last_fallback()[0] == fell
            for i in range(len(pairs)):
                f, o, fl = frames.view(i), outs.view(1).view(i), flags.view(1).view(i)
                pa.run_async(f.view(0).value, f.view(1).value, o.view(0).value, o.view(1).value, fl.view(0).value,
                             fl.view(1).value)
            pa.synchronize()
            got, want = outs.get(), flags.get()
            assert np.array_equal(got[0], got[1]) and np.array_equal(want[0], want[1]), (fell, rep)
        assert pa.learned_counts() == (2 * len(pairs), 2 * len(pairs) if fell else 0)
        assert ps.learned_counts() == pa.learned_counts()
