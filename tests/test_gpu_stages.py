"""Stage-level parity of the HIP kernels against the CPU oracle, through the C
ABI stage entry points.  Tolerances:
  * integer peak indices and the window grid bit-exact, except windows whose
    two largest correlation values tie within 1e-5 (counted, asserted rare);
  * u, v within 1e-3 px max-abs of the oracle's FP32 path;
  * validation flags identical except vectors whose median-residual ratio is
    within 1e-5 of the threshold.
"""

import numpy as np
import pytest

from oracle import piv_oracle as O

pytestmark = pytest.mark.gpu

gu = pytest.importorskip("gpu_util")


def _frames(shape, dtype, seed, shift=(2.35, -1.6)):
    from openpiv_gpu import synthetic
    if dtype == np.float32:
        rng = np.random.default_rng(seed)
        a, b = synthetic.make_pair(shape, synthetic.uniform_displacement(*shift), synthetic.ParticleModel(), seed)
        return (a.astype(np.float32) + rng.random(shape, dtype=np.float32),
                b.astype(np.float32) + rng.random(shape, dtype=np.float32))
    if dtype == np.uint16:
        model = synthetic.ParticleModel(i0_low=1000, i0_high=4095, noise_sigma=20, max_value=4095, dtype=np.uint16)
    else:
        model = synthetic.ParticleModel()
    return synthetic.make_pair(shape, synthetic.uniform_displacement(*shift), model, seed)


@pytest.mark.parametrize("ss,n_fft", [(8, 2), (16, 2), (32, 2), (64, 2), (24, 2), (48, 2), (16, 3), (32, 1)])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_correlate_matches_oracle(ss, n_fft, dtype):
    fsize = O.fft_size(ss, n_fft)
    shape = (max(6 * ss, 96), max(5 * ss, 80))
    fa, fb = _frames(shape, dtype, seed=ss + n_fft)
    geom = O.field_geometry(shape, ss, 0.5)
    ref = O.correlate_pass(fa, fb, geom, n_fft=n_fft, return_planes=True)
    got = gu.stage_correlate(fa, fb, geom, fsize, planes=True)
    planes_ref = ref["corr"]
    scale = np.abs(planes_ref).reshape(len(planes_ref), -1).max(axis=1)[:, None, None]
    err = np.abs(got["corr"] - planes_ref) / np.maximum(scale, 1e-30)
    assert err.max() < 2e-5, f"plane rel err {err.max()}"
    ties = gu.near_tie_mask(planes_ref).reshape(geom.field_shape)
    assert ties.mean() < 0.01
    ok = ~ties
    assert np.array_equal(got["i_peak"][ok], ref["i_peak"][ok])
    assert np.array_equal(got["j_peak"][ok], ref["j_peak"][ok])
    amb = gu.fit_sign_ambiguous(planes_ref, ref["i_peak"], ref["j_peak"]).reshape(geom.field_shape)
    assert amb.mean() < (0.05 if ss <= 8 else 0.02)  # 8 px windows: noisy planes
    ok &= ~amb
    assert np.abs(got["u"] - ref["u"])[ok].max() <= 1e-3
    assert np.abs(got["v"] - ref["v"])[ok].max() <= 1e-3


@pytest.mark.parametrize("method,code", [("gaussian", 0), ("parabolic", 1), ("centroid", 2)])
def test_correlate_subpixel_methods_and_predictor(method, code):
    shape = (256, 256)
    fa, fb = _frames(shape, np.uint8, seed=11)
    geom = O.field_geometry(shape, 32, 0.5)
    ref = O.correlate_pass(fa, fb, geom, subpixel_method=method, return_planes=True)
    rng = np.random.default_rng(0)
    up = rng.uniform(-3, 3, geom.field_shape).astype(np.float32)
    vpred = rng.uniform(-3, 3, geom.field_shape).astype(np.float32)
    got = gu.stage_correlate(fa, fb, geom, 64, up=up, vpred=vpred, subpixel=code)
    ok = ~gu.near_tie_mask(ref["corr"]).reshape(geom.field_shape)
    assert np.abs(got["u"] - (ref["u"] + up))[ok].max() <= 1e-3
    assert np.abs(got["v"] - (ref["v"] + vpred))[ok].max() <= 1e-3


def test_correlate_blank_window_gives_corner_peak():
    # all-constant windows correlate to exact zeros: the first occurrence
    # (index 0) wins and the border keeps the integer peak -> u = v = -F/2
    fa = np.full((64, 64), 7, np.uint8)
    geom = O.field_geometry(fa.shape, 16, 0.5)
    got = gu.stage_correlate(fa, fa, geom, 32)
    ref = O.correlate_pass(fa, fa, geom)
    assert np.array_equal(got["u"], ref["u"]) and np.all(got["u"] == -16)


@pytest.mark.parametrize("deforming_par", [0.5, 1.0, 0.0, 0.3])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_deform_matches_oracle(deforming_par, dtype):
    shape = (160, 200)
    fa, fb = _frames(shape, dtype, seed=3)
    geom = O.field_geometry(shape, 32, 0.5)
    rng = np.random.default_rng(1)
    up = rng.uniform(-4, 4, geom.field_shape)
    vpred = rng.uniform(-4, 4, geom.field_shape)
    ra, rb = O.deform_frames(fa, fb, geom, up, vpred, deforming_par)
    ga, gb = gu.stage_deform(fa, fb, geom, up.astype(np.float32), vpred.astype(np.float32), deforming_par)
    tol = 1e-3 * float(max(fa.max(), fb.max()))  # FP32 warp vs FP64-interpolated predictor
    assert np.abs(ga - ra).max() <= tol
    assert np.abs(gb - rb).max() <= tol


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("wd", [130, 195, 67])
def test_deform_narrow_last_wave_matches_oracle(dtype, wd):
    """Frame widths whose last wave holds one to three columns: lane k of a
    wave supplies row k's predictor cell to the others, so every lane forms
    the predictor before the columns beyond the frame leave."""
    shape = (96, wd)
    fa, fb = _frames(shape, dtype, seed=11)
    geom = O.field_geometry(shape, 32, 0.5)
    rng = np.random.default_rng(12)
    up = rng.uniform(-4, 4, geom.field_shape)
    vpred = rng.uniform(-4, 4, geom.field_shape)
    ra, rb = O.deform_frames(fa, fb, geom, up, vpred, 0.5)
    ga, gb = gu.stage_deform(fa, fb, geom, up.astype(np.float32), vpred.astype(np.float32), 0.5)
    tol = 1e-3 * float(max(fa.max(), fb.max()))  # FP32 warp vs FP64-interpolated predictor
    assert np.abs(ga - ra).max() <= tol
    assert np.abs(gb - rb).max() <= tol


@pytest.mark.parametrize("shape,win,amp,par", [((160, 200), 32, 4.0, 0.5), ((257, 130), 32, 30.0, 0.3),
                                              ((64, 66), 16, 60.0, 0.5), ((300, 512), 8, 12.0, 0.5),
                                              ((96, 34), 16, 3.0, 0.5)])
def test_deform_uint16_matches_float_frames(shape, win, amp, par):
    """uint16 frames against the same frames as float32: the integer tap
    loads and their conversion give bit-identical deformed frames, including
    large displacements, frame edges and widths whose last wave is nearly
    empty."""
    fa, fb = _frames(shape, np.uint16, seed=5)
    geom = O.field_geometry(shape, win, 0.5)
    rng = np.random.default_rng(7)
    up = rng.uniform(-amp, amp, geom.field_shape).astype(np.float32)
    vpred = rng.uniform(-amp, amp, geom.field_shape).astype(np.float32)
    wa, wb = gu.stage_deform(fa, fb, geom, up, vpred, par)
    ra, rb = gu.stage_deform(fa.astype(np.float32), fb.astype(np.float32), geom, up, vpred, par)
    assert np.array_equal(wa.view(np.uint32), ra.view(np.uint32))
    assert np.array_equal(wb.view(np.uint32), rb.view(np.uint32))


def test_grid_interp_matches_oracle():
    shape = (512, 640)
    g1 = O.field_geometry(shape, 64, 0.5)
    g2 = O.field_geometry(shape, 32, 0.5)
    rng = np.random.default_rng(2)
    u = rng.uniform(-5, 5, g1.field_shape)
    v = rng.uniform(-5, 5, g1.field_shape)
    ru, rv = O.interpolate_displacement(g1, g2, u, v)
    gu_, gv_ = gu.stage_grid_interp(g1, g2, u, v)
    assert np.abs(gu_ - ru).max() < 1e-5 and np.abs(gv_ - rv).max() < 1e-5


def _noisy_field(shape, seed, n_out):
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    u = 3 * np.sin(xx / 9.0) + 0.3 * rng.standard_normal(shape)
    v = 2 * np.cos(yy / 7.0) + 0.3 * rng.standard_normal(shape)
    idx = rng.choice(u.size, n_out, replace=False)
    u.flat[idx] += rng.uniform(-15, 15, n_out)
    v.flat[idx[: n_out // 2]] += rng.uniform(-15, 15, n_out // 2)
    return u.astype(np.float32), v.astype(np.float32)


@pytest.mark.parametrize("shape,n_out", [((31, 31), 20), ((127, 127), 300), ((255, 255), 2000), ((1, 9), 2),
                                         ((2, 2), 1)])
def test_validate_matches_oracle(shape, n_out):
    u, v = _noisy_field(shape, 4, n_out)
    ref = O.median_validation(u, v, 2.0)
    got, count = gu.stage_validate(u, v, 2.0)
    ratio = O.median_residual_ratio(u, v, 2.0)
    band = np.abs(ratio - 1.0) <= 1e-5
    assert np.array_equal(got[~band], ref[~band])
    assert count == int(got.sum())


@pytest.mark.parametrize("shape,n_out", [((31, 31), 20), ((127, 127), 400), ((63, 63), 1500)])
def test_spring_replace_matches_oracle(shape, n_out):
    u, v = _noisy_field(shape, 5, n_out)
    flags = O.median_validation(u, v, 2.0)
    ru, rv = O.spring_replacement(u.astype(np.float64), v.astype(np.float64), flags)
    gu_, gv_ = gu.stage_replace(u, v, flags)
    assert np.abs(gu_ - ru).max() < 1e-4 and np.abs(gv_ - rv).max() < 1e-4
    assert np.array_equal(gu_[~flags], u[~flags])


def _planted_clusters(shape, sizes, seed):
    """Flags with outlier clusters of the given sizes (random 4-connected
    blobs, corners and borders included), kept apart from each other."""
    rng = np.random.default_rng(seed)
    m, n = shape
    flags = np.zeros(shape, bool)
    taken = np.zeros(shape, bool)  # clusters and their 4-neighbours
    starts = [(0, 0), (m - 1, n - 1), (0, n // 2), (m // 2, 0)]
    for c, size in enumerate(sizes):
        for _ in range(200):
            i, j = starts[c] if c < len(starts) else (int(rng.integers(m)), int(rng.integers(n)))
            blob = [(i, j)]
            while len(blob) < size:
                bi, bj = blob[int(rng.integers(len(blob)))]
                di, dj = [(0, 1), (0, -1), (1, 0), (-1, 0)][int(rng.integers(4))]
                a, b = bi + di, bj + dj
                if 0 <= a < m and 0 <= b < n and (a, b) not in blob:
                    blob.append((a, b))
            if not any(taken[a, b] for a, b in blob):
                break
            starts = []
        for a, b in blob:
            flags[a, b] = True
        for a, b in blob:
            for da, db in ((0, 0), (0, 1), (0, -1), (1, 0), (-1, 0)):
                if 0 <= a + da < m and 0 <= b + db < n:
                    taken[a + da, b + db] = True
    return flags


@pytest.mark.parametrize("sizes", [
    [1, 2, 3, 4, 5, 6, 2, 3, 6, 5, 1, 4] * 4,   # every cluster solved directly, no CG
    [9, 2, 7, 3, 1, 6, 12, 4, 5, 2],            # larger clusters left to CG
    [1] * 30,                                   # isolated outliers only
])
def test_spring_replace_clusters_match_oracle(sizes):
    """Small outlier clusters are solved exactly per cluster and larger ones by
    CG started from those solutions; both must equal the sparse direct solve
    of the whole system (gpu_validation.py:336-371)."""
    u, v = _noisy_field((63, 63), 8, 0)
    flags = _planted_clusters((63, 63), sizes, 9)
    assert 0 < int(flags.sum()) <= 256
    u = np.where(flags, u + 12.0, u).astype(np.float32)
    ru, rv = O.spring_replacement(u.astype(np.float64), v.astype(np.float64), flags)
    gu_, gv_ = gu.stage_replace(u, v, flags)
    # tolerance: FP32 storage of the solution plus the CG stopping rule
    assert np.abs(gu_ - ru).max() < 2e-5 and np.abs(gv_ - rv).max() < 2e-5
    assert np.array_equal(gu_[~flags], u[~flags]) and np.array_equal(gv_[~flags], v[~flags])


def test_spring_replace_all_outliers_keeps_field():
    u, v = _noisy_field((9, 9), 6, 0)
    flags = np.ones((9, 9), bool)
    gu_, gv_ = gu.stage_replace(u, v, flags)
    assert np.array_equal(gu_, u) and np.array_equal(gv_, v)


@pytest.mark.parametrize("shape", [(63, 63), (127, 127), (17, 23)])
def test_smoothn_fixed_s(shape):
    u, v = _noisy_field(shape, 7, 0)
    rz = O.smoothn([u, v], s=2.5)
    gz_u, gz_v, p = gu.stage_smoothn(u, v, s=2.5)
    assert abs(p - np.log10(2.5)) < 1e-12
    assert np.abs(gz_u - rz[0]).max() < 5e-5 and np.abs(gz_v - rz[1]).max() < 5e-5


@pytest.mark.parametrize("shape", [(63, 63), (127, 127)])
def test_smoothn_auto_s(shape):
    u, v = _noisy_field(shape, 8, 0)
    (ru, rv), p_ref = O.smoothn([u, v], return_p=True)
    p_exact = O.gcv_exact_minimiser([u, v])
    gz_u, gz_v, p = gu.stage_smoothn(u, v)
    assert abs(p - p_exact) < 1e-6, (p, p_exact, p_ref)
    assert abs(p - p_ref) < 5e-3
    assert np.abs(gz_u - ru).max() < 1e-3 and np.abs(gz_v - rv).max() < 1e-3


@pytest.mark.parametrize("ss,ws", [(32, 20), (64, 32), (16, 8)])
def test_correlate_shrunk_interrogation_window(ss, ws):
    """First pass with shrink_ratio: frame A's interrogation window inside the
    search window (gpu_process.py:654-662, 1066-1081)."""
    shape = (6 * ss, 5 * ss)
    fa, fb = _frames(shape, np.uint8, seed=ss + ws)
    geom = O.field_geometry(shape, ss, 0.5)
    fsize = O.fft_size(ss)
    ref = O.correlate_pass(fa, fb, geom, return_planes=True, window_size=ws)
    got = gu.stage_correlate(fa, fb, geom, fsize, planes=True, window_size=ws)
    scale = np.abs(ref["corr"]).reshape(len(ref["corr"]), -1).max(axis=1)[:, None, None]
    assert (np.abs(got["corr"] - ref["corr"]) / scale).max() < 2e-5
    ok = ~gu.near_tie_mask(ref["corr"]).reshape(geom.field_shape)
    ok &= ~gu.fit_sign_ambiguous(ref["corr"], ref["i_peak"], ref["j_peak"]).reshape(geom.field_shape)
    assert np.array_equal(got["i_peak"][ok], ref["i_peak"][ok])
    assert np.abs(got["u"] - ref["u"])[ok].max() <= 1e-3 and np.abs(got["v"] - ref["v"])[ok].max() <= 1e-3


def test_shrink_ratio_end_to_end():
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (384, 384)
    fa, fb = _frames(shape, np.uint8, seed=77, shift=(5.5, -4.2))
    # fixed smoothing parameter: both sides smooth the first pass identically
    # (the GCV search itself is checked elsewhere)
    piv = PIVGPU(shape, 16, search_size_iters=(1, 1), shrink_ratio=0.6, median_tol=2, smoothing_par=0.8,
                 dtype_f="float32")
    u, v = piv.run_host(fa, fb)
    s = O.OracleSettings(min_search_size=16, search_size_iters=(1, 1), shrink_ratio=0.6, smoothing_par=0.8)
    r = O.piv(fa, fb, s, keep_passes=True)
    excl = gu.pass_exclusions(fa, fb, s, r["passes"])
    rep = gu.e2e_compare(u, v, piv.init_val_locations, r, excl, s, final_flags=piv.val_locations)
    assert rep["compared"] > 0.95 * u.size, rep
