"""The rest of validation and replacement on the GPU against the oracle:
signal-to-noise ratios (peak2peak / peak2mean / peak2energy,
gpu_process.py:1312-1365), the S2N / median / MAD / mean / RMS tests
(gpu_validation.py:84-164), median and mean replacement with iterations and
revalidation (gpu_process.py:788-834; gpu_validation.py:294-415), and a
multi-pass run with all of them.

Tolerances: S2N within 1e-5 relative of the oracle evaluated on the kernel's
own planes; against the oracle's FP32 planes 1e-4 for peak2peak and ~1 %
for the two mean methods, whose value sets move with the sign of near-zero
plane values (FFT rounding) -- windows whose second-peak candidates tie
within 1e-5 are left out; flags identical except vectors within 1e-4 (relative) of a test's
threshold; replaced vectors within 1e-5 px."""

import numpy as np
import pytest

from oracle import piv_oracle as O

pytestmark = pytest.mark.gpu
gu = pytest.importorskip("gpu_util")


def _frames(shape, seed, shift=(2.35, -1.6)):
    from openpiv_gpu import synthetic
    return synthetic.make_pair(shape, synthetic.uniform_displacement(*shift), synthetic.ParticleModel(), seed)


def _second_peak_tie(planes, ip, jp, size, rel=1e-5):
    out = np.zeros(len(planes), bool)
    F = planes.shape[1]
    for k in range(len(planes)):
        p = planes[k].copy()
        sh = np.arange(-size, size + 1)
        p[np.ix_((ip[k] + sh) % F, (jp[k] + sh) % F)] = -np.inf
        flat = np.sort(p.reshape(-1))
        out[k] = flat[-1] - flat[-2] <= rel * max(abs(flat[-1]), 1e-30)
    return out


@pytest.mark.parametrize("method,size", [("peak2peak", 2), ("peak2peak", 4), ("peak2mean", 2), ("peak2energy", 2)])
def test_s2n_matches_oracle(method, size):
    shape = (256, 320)
    fa, fb = _frames(shape, 31)
    geom = O.field_geometry(shape, 32, 0.5)
    got = gu.stage_correlate(fa, fb, geom, 64, planes=True, s2n_method=method, s2n_size=size)
    ip, jp = got["i_peak"].reshape(-1), got["j_peak"].reshape(-1)
    have = got["s2n"].reshape(-1)
    # exact semantics: the oracle's S2N of the very planes the kernel produced
    # (only the summation order of the two means differs)
    want = O.s2n_ratio(got["corr"], ip, jp, method, size)
    ok = ~gu.near_tie_mask(got["corr"])
    if method == "peak2peak":
        ok &= ~_second_peak_tie(got["corr"], ip, jp, size)
    assert ok.mean() > 0.97
    fin = ok & np.isfinite(want)
    assert np.all(np.abs(have[fin] - want[fin]) <= 1e-5 * np.abs(want[fin]))
    assert np.array_equal(np.isinf(have[ok]), np.isinf(want[ok]))
    # against the oracle's own FP32 planes: peak2peak to 1e-4; the two means
    # move with the sign of near-zero plane values (FFT rounding), ~1 %
    ref = O.correlate_pass(fa, fb, geom, return_planes=True)
    rp = O.s2n_ratio(ref["corr"], ref["i_peak"].reshape(-1), ref["j_peak"].reshape(-1), method, size)
    same = (ip == ref["i_peak"].reshape(-1)) & (jp == ref["j_peak"].reshape(-1)) & ok & np.isfinite(rp)
    if method == "peak2peak":
        same &= ~_second_peak_tie(ref["corr"], ip, jp, size)
    rel = np.abs(have[same] - rp[same]) / np.abs(rp[same])
    if method == "peak2peak":
        assert rel.max() <= 1e-4
    else:
        assert np.median(rel) <= 1e-2 and rel.max() <= 5e-2


def _field(seed, shape=(40, 48), n_out=60):
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    u = (2 * np.sin(xx / 7.0) + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    v = (1.5 * np.cos(yy / 5.0) + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    idx = rng.choice(u.size, n_out, replace=False)
    u.flat[idx] += rng.choice([-1, 1], n_out) * rng.uniform(3, 9, n_out)
    v.flat[idx[::2]] += rng.uniform(-8, 8, idx[::2].size)
    return u, v


def _band(u, v, s2n, tests, rel=1e-4, size=1):
    """Vectors within rel of any enabled test's threshold."""
    med_u, mad_u, mean_u = O.neighbour_stats(u, size)
    med_v, mad_v, mean_v = O.neighbour_stats(v, size)
    near = np.zeros(u.shape, bool)

    def close(x, t):
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.abs(x - t) <= rel * np.maximum(np.abs(t), 1e-12)

    for f, med, mad, mean in ((u, med_u, mad_u, mean_u), (v, med_v, mad_v, mean_v)):
        if tests.get("median_tol") is not None:
            near |= close(np.abs(f - med), tests["median_tol"])
        if tests.get("mad_tol") is not None:
            near |= close(np.abs(f - med), tests["mad_tol"] * mad)
        if tests.get("mean_tol") is not None:
            near |= close(np.abs(f - mean), tests["mean_tol"])
        if tests.get("rms_tol") is not None:
            near |= close(np.abs(f - f.mean(dtype=np.float64)), tests["rms_tol"] * f.std(dtype=np.float64))
    if tests.get("s2n_tol") is not None:
        near |= close(s2n, tests["s2n_tol"])
    return near


@pytest.mark.parametrize("tests", [
    dict(median_tol=2.0), dict(median_tol=None, mad_tol=3.0), dict(median_tol=None, mean_tol=1.5),
    dict(median_tol=None, rms_tol=2.0), dict(median_tol=None, s2n_tol=1.3),
    dict(median_tol=1.0, mad_tol=4.0, mean_tol=2.0, rms_tol=2.5, s2n_tol=1.2),
])
def test_validation_tests_match_oracle(tests):
    u, v = _field(3)
    rng = np.random.default_rng(4)
    s2n = rng.uniform(1.0, 2.0, u.shape).astype(np.float32)
    desc = gu.tests_desc(**tests)
    got, count = gu.stage_validate_all(u, v, desc, s2n=s2n if "s2n_tol" in tests else None)
    ref = O.validate(u, v, s2n if "s2n_tol" in tests else None, tests.get("s2n_tol"), tests.get("median_tol"),
                     tests.get("mad_tol"), tests.get("mean_tol"), tests.get("rms_tol"))
    ok = ~_band(u.astype(np.float64), v.astype(np.float64), s2n, tests)
    assert np.array_equal(got[ok], ref[ok]) and count == int(got.sum()) and got.sum() > 0
    # restricted to a subset
    only = rng.random(u.shape) < 0.3
    got_o, _ = gu.stage_validate_all(u, v, desc, s2n=s2n if "s2n_tol" in tests else None, only=only)
    assert not got_o[~only].any() and np.array_equal(got_o[ok & only], ref[ok & only])


@pytest.mark.parametrize("method", ["median", "mean", "spring"])
@pytest.mark.parametrize("n_iters,revalidate", [(1, False), (3, False), (2, True), (4, True)])
def test_replacement_loop_matches_oracle(method, n_iters, revalidate):
    u, v = _field(5)
    tests = dict(median_tol=1.0)
    flags = O.validate(u, v, median_tol=1.0)
    desc = gu.tests_desc(method=method, n_iters=n_iters, revalidate=revalidate, **tests)
    gu_, gv_, gflags = gu.stage_replace_all(u, v, flags, desc)
    fn = lambda a, b, only=None: O.validate(a, b, median_tol=1.0, only=only)
    ru, rv, rflags = O.replace_outliers(u, v, flags, method, n_iters, revalidate, fn)
    assert np.abs(gu_ - ru).max() <= 1e-4 and np.abs(gv_ - rv).max() <= 1e-4
    assert np.array_equal(gflags, rflags)
    assert np.array_equal(gu_[~flags], u[~flags])


@pytest.mark.parametrize("size", [2, 3, 8])
@pytest.mark.parametrize("tests", [dict(median_tol=2.0), dict(median_tol=None, mad_tol=3.0),
                                   dict(median_tol=1.0, mad_tol=4.0, mean_tol=2.0)])
def test_wide_validation_matches_oracle(size, tests):
    """validation_size 2, 3 and the largest, 8: (2 size + 1)^2 neighbourhoods
    (gpu_validation.py:73-76, 194-205), NaN vectors left out."""
    u, v = _field(13)
    u[7, 9] = np.nan
    v[20, 3] = np.nan
    desc = gu.tests_desc(validation_size=size, **tests)
    got, count = gu.stage_validate_all(u, v, desc)
    ref = O.validate(u, v, None, None, tests.get("median_tol"), tests.get("mad_tol"), tests.get("mean_tol"),
                     None, size=size)
    ok = ~_band(u.astype(np.float64), v.astype(np.float64), None, tests, size=size)
    assert ok.mean() > 0.98
    assert np.array_equal(got[ok], ref[ok]) and count == int(got.sum()) and got.sum() > 0


@pytest.mark.parametrize("method", ["median", "mean", "spring"])
@pytest.mark.parametrize("size", [2, 3])
@pytest.mark.parametrize("n_iters,revalidate", [(1, False), (3, False), (2, True)])
def test_wide_replacement_matches_oracle(method, size, n_iters, revalidate):
    """replacing_size 2 and 3 (gpu_validation.py:272-290, 336-398): median
    and mean over (2 size + 1)^2 neighbourhoods, spring links size apart;
    revalidation with validation_size = size."""
    u, v = _field(17)
    flags = O.validate(u, v, median_tol=1.0, size=size)
    assert not O.spring_singular(flags, size).any()
    desc = gu.tests_desc(method=method, n_iters=n_iters, revalidate=revalidate, median_tol=1.0,
                         validation_size=size, replacing_size=size)
    gu_, gv_, gflags = gu.stage_replace_all(u, v, flags, desc)
    fn = lambda a, b, only=None: O.validate(a, b, median_tol=1.0, size=size, only=only)
    ru, rv, rflags = O.replace_outliers(u, v, flags, method, n_iters, revalidate, fn, size)
    assert np.abs(gu_ - ru).max() <= 1e-4 and np.abs(gv_ - rv).max() <= 1e-4
    assert np.array_equal(gflags, rflags)
    assert np.array_equal(gu_[~flags], u[~flags])


def test_wide_spring_singular_class_keeps_its_values():
    """A residue class (rows and columns mod size) flagged entirely is a
    singular block of the size-2 spring system: its vectors stay as they
    were, the other outliers are solved."""
    rng = np.random.default_rng(19)
    u = rng.standard_normal((12, 14)).astype(np.float32)
    v = rng.standard_normal((12, 14)).astype(np.float32)
    flags = np.zeros((12, 14), bool)
    flags[1::2, 0::2] = True
    flags[6, 7] = flags[2, 3] = True
    assert O.spring_singular(flags, 2).sum() == 6 * 7
    desc = gu.tests_desc(method="spring", replacing_size=2)
    gu_, gv_, _ = gu.stage_replace_all(u, v, flags, desc)
    ru, rv, _ = O.replace_outliers(u, v, flags, "spring", 1, size=2)
    assert np.array_equal(gu_[1::2, 0::2], u[1::2, 0::2])
    assert np.abs(gu_ - ru).max() <= 1e-5 and np.abs(gv_ - rv).max() <= 1e-5


def test_unresolved_block_is_reset_on_gpu():
    u = np.ones((10, 10), np.float32)
    v = np.ones((10, 10), np.float32)
    flags = np.zeros((10, 10), bool)
    flags[2:8, 2:8] = True
    u[flags] = 7.0
    for method in ("median", "mean"):
        gu_, gv_, _ = gu.stage_replace_all(u, v, flags, gu.tests_desc(method=method))
        ru, rv, _ = O.replace_outliers(u, v, flags, method, 1)
        assert np.array_equal(gu_, ru.astype(np.float32)) and not np.isnan(gu_).any()


@pytest.mark.parametrize("s2n_method", ["peak2energy", "peak2peak"])
def test_multipass_with_every_test_matches_oracle(s2n_method):
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (320, 384)
    fa, fb = _frames(shape, 41, shift=(3.2, 1.7))
    kw = dict(search_size_iters=(1, 1), overlap_ratio=0.5, s2n_method=s2n_method, s2n_tol=1.05, median_tol=2.0,
              mad_tol=6.0, mean_tol=3.0, replacing_method="median", num_replacing_iters=2, revalidate=True)
    piv = PIVGPU(shape, 16, dtype_f="float32", **kw)
    u, v = piv.run_host(fa, fb)
    s = O.OracleSettings(min_search_size=16, **kw)
    r = O.piv(fa, fb, s, keep_passes=True)
    # the mean-based S2N moves with the sign of near-zero plane values: a
    # few % (see test_s2n_matches_oracle); peak2peak to 1e-4
    excl = gu.pass_exclusions(fa, fb, s, r["passes"], s2n_rel=5e-2 if s2n_method == "peak2energy" else 1e-4)
    rep = gu.e2e_compare(u, v, piv.init_val_locations, r, excl, s, final_flags=piv.val_locations)
    assert rep["compared"] > 0.9 * u.size, rep
    for k, st in enumerate(r["passes"]):
        assert piv.outlier_counts(k)[0] == int(st.val.sum()), k


@pytest.mark.parametrize("method,vsize,rsize", [("spring", 2, 2), ("median", 3, 2), ("mean", 1, 3)])
def test_multipass_wide_kernels_match_oracle(method, vsize, rsize):
    """A three-pass run with validation_size / replacing_size above 1 in
    every pass (gpu_process.py:747-780)."""
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (320, 384)
    fa, fb = _frames(shape, 47, shift=(3.2, 1.7))
    fb = fb.copy()
    rng = np.random.default_rng(47)
    for _ in range(5):
        r0, c0 = rng.integers(20, 260, 2)
        fb[r0:r0 + 48, c0:c0 + 48] = rng.integers(0, 255, (48, 48))
    kw = dict(search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=1.0, replacing_method=method,
              num_replacing_iters=2, validation_size=vsize, replacing_size=rsize, smoothing_par=1.5)
    piv = PIVGPU(shape, 8, dtype_f="float32", **kw)
    u, v = piv.run_host(fa, fb)
    s = O.OracleSettings(min_search_size=8, **kw)
    r = O.piv(fa, fb, s, keep_passes=True)
    excl = gu.pass_exclusions(fa, fb, s, r["passes"])
    rep = gu.e2e_compare(u, v, piv.init_val_locations, r, excl, s, final_flags=piv.val_locations)
    # ~8 % of the last field are outliers; a spring cluster touched by one of
    # the few fit-ambiguous windows is excused whole, so fewer vectors remain
    assert rep["compared"] > (0.85 if method == "spring" else 0.9) * u.size, rep
    for k, st in enumerate(r["passes"]):
        assert piv.outlier_counts(k)[0] == int(st.val.sum()), k
    assert all(int(st.val.sum()) > 0 for st in r["passes"])


@pytest.mark.parametrize("graph", [False, True])
def test_multipass_spring_revalidate_counts_per_pass(graph):
    """Spring replacement with revalidation and several iterations in every
    pass: each pass's revalidation counters start from zero (a count summed
    over passes would reach the spring kernel's all-outlier gate and skip
    the replacement), on the eager run and on graph replays."""
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (320, 384)
    fa, fb = _frames(shape, 43, shift=(3.2, 1.7))
    fb = fb.copy()
    rng = np.random.default_rng(43)
    for _ in range(5):
        r0, c0 = rng.integers(30, 260, 2)
        fb[r0:r0 + 30, c0:c0 + 30] = rng.integers(0, 255, (30, 30))
    kw = dict(search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=1.0, replacing_method="spring",
              num_replacing_iters=3, revalidate=True, smoothing_par=1.5)
    piv = PIVGPU(shape, 8, dtype_f="float32", use_graph=graph, **kw)
    s = O.OracleSettings(min_search_size=8, **kw)
    r = O.piv(fa, fb, s, keep_passes=True)
    excl = gu.pass_exclusions(fa, fb, s, r["passes"])
    for call in range(3 if graph else 1):
        u, v = piv.run_host(fa, fb)
        rep = gu.e2e_compare(u, v, piv.init_val_locations, r, excl, s, final_flags=piv.val_locations)
        assert rep["compared"] > 0.9 * u.size, (call, rep)
        for k, st in enumerate(r["passes"]):
            assert piv.outlier_counts(k)[0] == int(st.val.sum()), (call, k)
        assert piv.outlier_counts(len(r["passes"]) - 1)[1] == int(r["outliers"].sum()), call
    assert all(int(st.val.sum()) > 0 for st in r["passes"])


@pytest.mark.parametrize("dtype_f,method", [("float32", "peak2peak"), ("float32", "peak2energy"),
                                            ("float64", "peak2mean")])
def test_s2n_ratio_property_matches_oracle(dtype_f, method):
    """PIVGPU.s2n_ratio: the last pass's S2N (gpu_process.py:864-867,
    1312-1376), recomputed from the frames that pass read; with a mask, 0
    inside it."""
    from openpiv_gpu import synthetic
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (256, 320)
    a, b = synthetic.make_pair(shape, synthetic.sine_displacement(2.0, 1.5, 128.0), synthetic.ParticleModel(), 41)
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    mask = (yy - 120) ** 2 + (xx - 150) ** 2 < 30 ** 2
    piv = PIVGPU(shape, 16, search_size_iters=(1, 1), s2n_method=method, mask=mask, smoothing_par=2.0,
                 dtype_f=dtype_f)
    piv(a, b)
    got = piv.s2n_ratio
    st = O.OracleSettings(min_search_size=16, search_size_iters=(1, 1), s2n_method=method, s2n_tol=0.0, mask=mask,
                          smoothing_par=2.0)
    ref = O.piv(a, b, st, mode="fp64" if dtype_f == "float64" else "faithful32", keep_passes=True)
    want = ref["passes"][-1].corr["s2n"]
    fm = O.field_mask(ref["passes"][-1].geom, mask)
    assert got.dtype == np.dtype(dtype_f) and got.shape == want.shape
    assert np.all(got[fm] == 0)
    fin = np.isfinite(want) & ~fm
    assert np.array_equal(np.isfinite(got) | fm, np.isfinite(want) | fm)
    rel = np.abs(got[fin] - want[fin]) / np.abs(want[fin])
    if method == "peak2peak":
        tol = 1e-7 if dtype_f == "float64" else 1e-3
    else:
        # the mean-based ratios count the positive plane values, and with
        # zero padding most of the plane is exactly 0 in exact arithmetic:
        # which of those come out positive is FFT rounding noise, in the
        # reference as here, so only the order of magnitude is reproducible
        tol = 0.2
    assert np.quantile(rel, 0.98) <= tol, np.quantile(rel, 0.98)
