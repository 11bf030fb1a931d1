"""Device particle-image renderer (piv_synth_pair) against the host model of
openpiv_gpu.synthetic: same blobs for the same particles, the host
displacement functions at the particle positions, run-to-run determinism,
and a known shift recovered by the PIV path."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gen(shape, model):
    from openpiv_gpu import synthetic
    return synthetic.DevicePairGenerator(shape, model)


@pytest.mark.parametrize("kind", ["uniform", "sine", "vortices"])
def test_blobs_match_host_render(kind):
    from openpiv_gpu import _native, synthetic as S
    shape = (192, 256)
    model = S.ParticleModel(noise_sigma=0.0, dtype=np.float32, max_value=1e9)
    disp = {"uniform": S.uniform_displacement(2.35, -1.6), "sine": S.sine_displacement(4.0, 3.0, 64.0),
            "vortices": S.vortex_field(4001, shape)}[kind]
    gen = _gen(shape, model)
    margin, n = S.particle_budget(shape, disp, model)
    parts = _native.DeviceBuffer((n, 5), np.float32)
    fa, fb = gen.render(disp, seed=77, particles=parts)
    p = parts.get().astype(np.float64)
    x, y, i0 = p[:, 0], p[:, 1], p[:, 2]
    assert x.min() >= -margin and x.max() < shape[1] + margin and y.min() >= -margin and y.max() < shape[0] + margin
    assert i0.min() >= model.i0_low and i0.max() <= model.i0_high
    du, dv = disp(x, y)
    assert np.abs(p[:, 3] - du).max() < 1e-4 * max(1.0, np.abs(du).max())
    assert np.abs(p[:, 4] - dv).max() < 1e-4 * max(1.0, np.abs(dv).max())
    # host render of the same particles (FP32-stored positions: allow the
    # rounding of a few pixels to land on the other side of .5)
    ra = np.clip(np.rint(S._render(*shape, x - p[:, 3] / 2, y - p[:, 4] / 2, i0, model.sigma)), 0, model.max_value)
    rb = np.clip(np.rint(S._render(*shape, x + p[:, 3] / 2, y + p[:, 4] / 2, i0, model.sigma)), 0, model.max_value)
    for got, ref in ((fa.get(), ra), (fb.get(), rb)):
        diff = np.abs(got - ref)
        assert diff.max() <= 1.0 and (diff > 0).mean() < 5e-3


def test_render_is_deterministic_and_noise_matches_model():
    from openpiv_gpu import synthetic as S
    shape = (256, 256)
    gen = _gen(shape, S.MODEL_12BIT)
    disp = S.vortex_field(4002, shape)
    a1, b1 = gen.render(disp, seed=5)
    a2, b2 = gen.render(disp, seed=5)
    assert np.array_equal(a1.get(), a2.get()) and np.array_equal(b1.get(), b2.get())
    a3, _ = gen.render(disp, seed=6)
    assert not np.array_equal(a1.get(), a3.get())
    # background pixels (no blob) carry the additive noise only
    flat = gen.render(S.uniform_displacement(0, 0), seed=9)[0].get().astype(np.float64)
    host_a, _ = S.make_pair(shape, S.uniform_displacement(0, 0), S.MODEL_12BIT, 9)
    assert abs(flat.mean() - host_a.mean()) < 0.1 * host_a.mean()
    assert flat.dtype == np.float64 and a1.dtype == np.uint16


def test_gpu_rendered_shift_is_recovered():
    from openpiv_gpu import synthetic as S
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (512, 512)
    gen = _gen(shape, S.ParticleModel())
    fa, fb = gen.render(S.uniform_displacement(2.35, -1.6), seed=1)
    piv = PIVGPU(shape, 32, overlap_ratio=0.5, median_tol=2, dtype_f="float32")
    u, v = piv.run_host(fa.get(), fb.get())
    assert np.sqrt(np.mean((u - 2.35) ** 2)) < 0.05 and np.sqrt(np.mean((v + 1.6) ** 2)) < 0.05


def test_plan_reads_device_frames_after_null_stream_producers():
    """Frames rendered on the null stream feed piv_run directly: the plan's
    own stream must wait for them (first, eager run and graph replays)."""
    from openpiv_gpu import _native, synthetic as S
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (1024, 1024)
    gen = _gen(shape, S.MODEL_12BIT)
    piv = PIVGPU(shape, 16, search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float32")
    plan = piv._get_plan(_native.PIV_U16)
    m, n = piv.piv_fields[-1].field_shape
    fa, fb = gen.new_frames()
    for seed in (4100, 4101, 4102):
        gen.render(S.vortex_field(seed, shape), seed, fa, fb)
        u = np.empty((m, n), np.float32)
        v = np.empty((m, n), np.float32)
        plan.run(fa.value, fb.value, True, u.ctypes.data, v.ctypes.data, None, None, False)
        uh, vh = piv.run_host(fa.get(), fb.get())
        assert np.array_equal(u, uh) and np.array_equal(v, vh), seed
