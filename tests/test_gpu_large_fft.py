"""FFT sizes above 128 on the GPU against the oracle: the reference pads
each window to 2^ceil(log2(search_size * n_fft)) points whatever the size
(gpu_process.py:1099-1109), so 128 px windows (F = 256), 48 / 64 px windows
with n_fft = 3 (F = 256), 256 px windows (F = 512) and 512 px windows
(F = 1024) all correlate.  Planes, integer peaks, fits, S2N and the shrunk
interrogation window in FP32 and FP64, and multi-pass runs whose first pass
is at F = 256.

Tolerances as for the small sizes: planes within 2e-5 (FP32) / 1e-12 (FP64)
of the plane maximum (the FFT's rounding grows with log F), integer peaks
identical outside near-tie windows, u / v within 1e-3 px (1e-9 in FP64)
outside near-ties and fit-sign-ambiguous windows."""

import numpy as np
import pytest

from oracle import piv_oracle as O

pytestmark = pytest.mark.gpu
gu = pytest.importorskip("gpu_util")


def _frames(shape, dtype, seed, shift=(3.4, -2.2)):
    from openpiv_gpu import synthetic
    if dtype == np.uint16:
        a, b, _ = synthetic.cfg_2048_uint16(seed, shape=shape)
        return a, b
    a, b = synthetic.make_pair(shape, synthetic.uniform_displacement(*shift), synthetic.ParticleModel(), seed)
    return a.astype(dtype), b.astype(dtype)


def _check(got, ref, shape, plane_tol=2e-5, uv_tol=1e-3, tie_rel=1e-5, fit_rel=1e-5):
    planes_ref = ref["corr"]
    scale = np.abs(planes_ref).reshape(len(planes_ref), -1).max(axis=1)[:, None, None]
    err = np.abs(got["corr"] - planes_ref) / np.maximum(scale, 1e-30)
    assert err.max() < plane_tol, f"plane rel err {err.max()}"
    ok = ~gu.near_tie_mask(planes_ref, rel=tie_rel).reshape(shape)
    assert ok.mean() > 0.95
    assert np.array_equal(got["i_peak"][ok], ref["i_peak"][ok])
    assert np.array_equal(got["j_peak"][ok], ref["j_peak"][ok])
    ok &= ~gu.fit_sign_ambiguous(planes_ref, ref["i_peak"], ref["j_peak"], rel=fit_rel).reshape(shape)
    assert ok.mean() > 0.9
    assert np.abs(got["u"] - ref["u"])[ok].max() <= uv_tol
    assert np.abs(got["v"] - ref["v"])[ok].max() <= uv_tol


@pytest.mark.parametrize("ss,n_fft,dtype", [
    (128, 2, np.uint8), (128, 2, np.uint16), (128, 2, np.float32),
    (48, 3, np.uint8), (64, 3, np.uint16), (256, 1, np.uint8),
    (256, 2, np.uint8), (160, 3, np.float32), (512, 2, np.uint8),
])
def test_large_fft_correlate_matches_oracle(ss, n_fft, dtype):
    fsize = O.fft_size(ss, n_fft)
    assert fsize > 128
    shape = (max(5 * ss, 256), max(4 * ss, 256))
    fa, fb = _frames(shape, dtype, seed=ss + n_fft)
    geom = O.field_geometry(shape, ss, 0.5)
    print(f"ss={ss} n_fft={n_fft} F={fsize} dtype={np.dtype(dtype).name} windows={geom.field_shape}", flush=True)
    ref = O.correlate_pass(fa, fb, geom, n_fft=n_fft, return_planes=True)
    got = gu.stage_correlate(fa, fb, geom, fsize, planes=True)
    _check(got, ref, geom.field_shape)


@pytest.mark.parametrize("method,code", [("parabolic", 1), ("centroid", 2)])
def test_large_fft_other_fits_and_predictor(method, code):
    shape = (640, 512)
    fa, fb = _frames(shape, np.uint8, seed=5)
    geom = O.field_geometry(shape, 128, 0.5)
    ref = O.correlate_pass(fa, fb, geom, subpixel_method=method, return_planes=True)
    rng = np.random.default_rng(1)
    up = rng.uniform(-3, 3, geom.field_shape).astype(np.float32)
    vpred = rng.uniform(-3, 3, geom.field_shape).astype(np.float32)
    got = gu.stage_correlate(fa, fb, geom, 256, up=up, vpred=vpred, subpixel=code)
    ok = ~gu.near_tie_mask(ref["corr"]).reshape(geom.field_shape)
    assert np.abs(got["u"] - (ref["u"] + up))[ok].max() <= 1e-3
    assert np.abs(got["v"] - (ref["v"] + vpred))[ok].max() <= 1e-3


@pytest.mark.parametrize("method,size", [("peak2peak", 2), ("peak2mean", 2), ("peak2energy", 3)])
def test_large_fft_s2n_matches_oracle_on_the_kernels_planes(method, size):
    shape = (640, 512)
    fa, fb = _frames(shape, np.uint8, seed=7)
    geom = O.field_geometry(shape, 128, 0.5)
    got = gu.stage_correlate(fa, fb, geom, 256, planes=True, s2n_method=method, s2n_size=size)
    ip, jp = got["i_peak"].reshape(-1), got["j_peak"].reshape(-1)
    want = O.s2n_ratio(got["corr"], ip, jp, method, size)
    have = got["s2n"].reshape(-1)
    fin = np.isfinite(want)
    assert fin.mean() > 0.9
    assert np.all(np.abs(have[fin] - want[fin]) <= 1e-5 * np.abs(want[fin]))
    assert np.array_equal(np.isinf(have), np.isinf(want))


def test_large_fft_interrogation_window_inside_search_window():
    """shrink_ratio: a 96 px interrogation window in 128 px search windows
    (gpu_process.py:1066-1083, 1194-1199), F = 256."""
    shape = (640, 512)
    fa, fb = _frames(shape, np.float32, seed=9)
    geom = O.field_geometry(shape, 128, 0.5)
    ref = O.correlate_pass(fa, fb, geom, return_planes=True, window_size=96)
    got = gu.stage_correlate(fa, fb, geom, 256, planes=True, window_size=96)
    _check(got, ref, geom.field_shape)


@pytest.mark.parametrize("ss,n_fft", [(128, 2), (64, 3)])
def test_large_fft_fp64_matches_fp64_oracle(ss, n_fft):
    shape = (640, 512)
    fa, fb = _frames(shape, np.uint16, seed=11)
    geom = O.field_geometry(shape, ss, 0.5)
    fsize = O.fft_size(ss, n_fft)
    got = gu.stage_correlate64(fa, fb, geom, fsize)
    ref = O.correlate_pass(np.asarray(fa, np.float64), np.asarray(fb, np.float64), geom, n_fft=n_fft,
                           dtype=np.float64, return_planes=True)
    _check(got, ref, geom.field_shape, plane_tol=1e-12, uv_tol=1e-9, tie_rel=1e-12, fit_rel=1e-10)


@pytest.mark.parametrize("n_fft,min_ss,iters", [(2, 32, (1, 1, 1)), (3, 32, (1, 1))])
def test_multipass_from_fft_256_matches_oracle(n_fft, min_ss, iters):
    """A multi-pass run whose first pass correlates at F = 256 (128 px
    windows with n_fft 2, or 64 px with n_fft 3), then deformation passes at
    smaller sizes."""
    from openpiv_gpu.gpu_process import PIVGPU
    from openpiv_gpu import synthetic
    shape = (768, 640)
    fa, fb = synthetic.make_pair(shape, synthetic.sine_displacement(3.0, 2.0, 200.0), synthetic.ParticleModel(), 13)
    kw = dict(search_size_iters=iters, overlap_ratio=0.5, n_fft=n_fft, smoothing_par=1.0)
    piv = PIVGPU(shape, min_ss, dtype_f="float32", **kw)
    u, v = piv.run_host(fa, fb)
    s = O.OracleSettings(min_search_size=min_ss, **kw)
    r = O.piv(fa, fb, s, keep_passes=True)
    assert O.fft_size(r["passes"][0].geom.search_size, n_fft) == 256
    excl = gu.pass_exclusions(fa, fb, s, r["passes"])
    rep = gu.e2e_compare(u, v, piv.init_val_locations, r, excl, s, final_flags=piv.val_locations)
    assert rep["compared"] > 0.9 * u.size, rep
