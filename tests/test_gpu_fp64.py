"""FP64 processing (dtype_f = "float64", gpu_process.py:552-562 with
USE_LONG kernels) on the GPU against the oracle's FP64 mode: the correlation
stage (planes, peaks, fits, S2N, interrogation window inside the search
window), single- and multi-pass runs end to end, masks, and every
validation test."""


import numpy as np
import pytest

from oracle import piv_oracle as O

pytestmark = pytest.mark.gpu
gu = pytest.importorskip("gpu_util")


_stage64 = gu.stage_correlate64


def _pair(shape, dtype, seed):
    from openpiv_gpu import synthetic
    if dtype == np.uint16:
        a, b, _ = synthetic.cfg_2048_uint16(seed, shape=shape)
    else:
        a, b = synthetic.make_pair(shape, synthetic.sine_displacement(3.0, 2.0, 96.0), synthetic.ParticleModel(), seed)
    return a.astype(dtype), b.astype(dtype)


@pytest.mark.parametrize("ss,dtype", [(8, np.uint8), (16, np.uint16), (32, np.float32), (64, np.float64)])
def test_correlate64_planes_and_peaks_match_fp64_oracle(ss, dtype):
    shape = (256, 320) if ss < 64 else (384, 448)
    a, b = _pair(shape, dtype, 3 + ss)
    g = O.field_geometry(shape, ss, 0.5)
    F = O.fft_size(ss)
    got = _stage64(a, b, g, F)
    ref = O.correlate_pass(np.asarray(a, np.float64), np.asarray(b, np.float64), g, dtype=np.float64,
                           return_planes=True)
    scale = np.abs(ref["corr"]).max(axis=(1, 2), keepdims=True)
    assert np.abs(got["corr"] - ref["corr"]).max() <= 1e-12 * scale.max()
    ties = gu.near_tie_mask(ref["corr"], rel=1e-12).reshape(g.field_shape)
    ok = ~ties
    assert np.array_equal(got["i_peak"][ok], ref["i_peak"][ok]) and np.array_equal(got["j_peak"][ok],
                                                                                   ref["j_peak"][ok])
    # a fit sample within rounding of 0 decides by its sign whether the fit runs
    amb = gu.fit_sign_ambiguous(ref["corr"], ref["i_peak"], ref["j_peak"], rel=1e-10).reshape(g.field_shape)
    ok &= ~amb
    assert amb.sum() <= max(2, amb.size // 20)
    assert np.abs(got["u"] - ref["u"])[ok].max() <= 1e-9 and np.abs(got["v"] - ref["v"])[ok].max() <= 1e-9


@pytest.mark.parametrize("subpixel", [1, 2])
def test_correlate64_other_fits_and_predictor(subpixel):
    a, b = _pair((256, 256), np.uint8, 9)
    g = O.field_geometry((256, 256), 32, 0.5)
    rng = np.random.default_rng(1)
    up = rng.uniform(-1, 1, g.field_shape)
    vpd = rng.uniform(-1, 1, g.field_shape)
    got = _stage64(a, b, g, 64, up=up, vpred=vpd, subpixel=subpixel)
    ref = O.correlate_pass(np.asarray(a, np.float64), np.asarray(b, np.float64), g, dtype=np.float64,
                           subpixel_method={1: "parabolic", 2: "centroid"}[subpixel])
    assert np.abs(got["u"] - (ref["u"] + up)).max() <= 1e-9 and np.abs(got["v"] - (ref["v"] + vpd)).max() <= 1e-9


@pytest.mark.parametrize("method,size", [("peak2peak", 2), ("peak2peak", 1), ("peak2mean", 2), ("peak2energy", 2)])
def test_correlate64_s2n_matches_oracle(method, size):
    a, b = _pair((256, 256), np.uint8, 11)
    g = O.field_geometry((256, 256), 32, 0.5)
    got = _stage64(a, b, g, 64, s2n_method=method, s2n_size=size)
    s = O.s2n_ratio(got["corr"], got["i_peak"], got["j_peak"], method, size).reshape(g.field_shape)
    fin = np.isfinite(s)
    assert np.array_equal(fin, np.isfinite(got["s2n"]))
    assert np.allclose(got["s2n"][fin], s[fin], rtol=1e-12, atol=0)


def test_correlate64_interrogation_window_inside_search_window():
    a, b = _pair((256, 256), np.uint16, 12)
    g = O.field_geometry((256, 256), 32, 0.5)
    got = _stage64(a, b, g, 64, window_size=16)
    ref = O.correlate_pass(np.asarray(a, np.float64), np.asarray(b, np.float64), g, dtype=np.float64,
                           return_planes=True, window_size=16)
    assert np.abs(got["corr"] - ref["corr"]).max() <= 1e-12 * np.abs(ref["corr"]).max()


def test_fp64_single_pass_512_uint8_matches_oracle():
    from openpiv_gpu import synthetic
    from openpiv_gpu.gpu_process import piv_gpu
    a, b, _ = synthetic.cfg_512_uint8(1)
    piv = piv_gpu((512, 512), 32, overlap_ratio=0.5, median_tol=2, dtype_f="float64")
    u, v = piv(a, b)
    ref = O.piv(a, b, O.OracleSettings(min_search_size=32), mode="fp64")
    assert u.dtype == np.float64 and u.shape == (31, 31)
    assert np.abs(u - ref["u"]).max() <= 1e-9 and np.abs(v - ref["v"]).max() <= 1e-9
    assert np.array_equal(piv.val_locations, ref["val_locations"])
    assert np.array_equal(piv.outliers, ref["outliers"])


@pytest.fixture(scope="module")
def fp64_runs():
    from openpiv_gpu import synthetic
    from openpiv_gpu.gpu_process import PIVGPU
    a, b, disp = synthetic.cfg_2048_uint16(2, shape=(1024, 1024))
    piv = PIVGPU((1024, 1024), 16, search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float64")
    runs = []
    for _ in range(3):  # eager, graph capture + replay, replay
        u, v = piv(a, b)
        runs.append((u.get(), v.get(), piv.init_val_locations.get().copy(),
                     [piv.smoothing_log10_s(k) for k in range(2)]))
    ref = O.piv(a, b, O.OracleSettings(min_search_size=16, search_size_iters=(1, 1, 1)), mode="fp64",
                keep_passes=True)
    return runs, ref


def test_fp64_three_pass_matches_oracle(fp64_runs):
    runs, ref = fp64_runs
    u, v, flags, ps = runs[0]
    assert u.dtype == np.float64 and u.shape == (127, 127)
    for k in range(2):
        assert abs(ps[k] - ref["passes"][k].smooth_p) < 5e-3, (k, ps[k], ref["passes"][k].smooth_p)
    assert np.abs(u - ref["u"]).max() <= 1e-3 and np.abs(v - ref["v"]).max() <= 1e-3
    assert np.array_equal(flags.astype(bool), ref["val_locations"])


def test_fp64_graph_replay_is_bitwise_stable(fp64_runs):
    runs, _ = fp64_runs
    for u, v, f, _ in runs[1:]:
        assert np.array_equal(u, runs[0][0]) and np.array_equal(v, runs[0][1]) and np.array_equal(f, runs[0][2])


def test_fp64_every_test_mask_and_scaling_match_oracle():
    from openpiv_gpu import synthetic
    from openpiv_gpu.gpu_process import piv_gpu
    shape = (320, 384)
    a, b = synthetic.make_pair(shape, synthetic.sine_displacement(2.0, 1.5, 160.0), synthetic.ParticleModel(), 31)
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    mask = (yy - 150) ** 2 + (xx - 190) ** 2 < 45 ** 2
    kw = dict(search_size_iters=(1, 2), overlap_ratio=0.5, median_tol=2.0, s2n_tol=1.05, mad_tol=6.0, rms_tol=5.0,
              replacing_method="median", num_replacing_iters=2, revalidate=True, smoothing_par=2.0, dt=0.3,
              scaling_par=1.7, mask=mask)
    piv = piv_gpu(shape, 16, dtype_f="float64", **kw)
    u, v = piv(a, b)
    st = O.OracleSettings(min_search_size=16, search_size_iters=(1, 2), median_tol=2.0, s2n_tol=1.05, mad_tol=6.0,
                          rms_tol=5.0, replacing_method="median", num_replacing_iters=2, revalidate=True,
                          smoothing_par=2.0, dt=0.3, scaling_par=1.7, mask=mask)
    ref = O.piv(a, b, st, mode="fp64")
    fm = np.isnan(ref["u"])
    assert np.array_equal(np.isnan(u), fm)
    assert np.abs(u - ref["u"])[~fm].max() <= 1e-6 and np.abs(v - ref["v"])[~fm].max() <= 1e-6
    assert np.array_equal(piv.val_locations, ref["val_locations"])
    assert np.array_equal(piv.outliers, ref["outliers"])
