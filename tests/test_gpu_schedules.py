"""FP32 plans with the reference's output scaling and repeated window sizes,
against the CPU oracle, and plan state after a rejected mask.

* ``dt`` / ``scaling_par`` != 1: ``u * scaling_par / dt`` and coordinates
  times ``scaling_par`` (gpu_process.py:642-643, 847-857);
* ``search_size_iters=(1, 2, 2)``, the tutorials' schedule: passes of
  64, 32, 32, 16, 16 px, where a repeated size takes the previous field as
  its predictor without interpolation (gpu_process.py:620-625, 675-680,
  730-742).

Tolerances: u, v within 1e-3 px (times scaling_par / dt) except the vectors
gpu_util.e2e_compare excuses by name (near-tie windows, fit-ambiguous
windows, threshold-band vectors and their replacement reach); flags
identical outside them."""

import numpy as np
import pytest

from oracle import piv_oracle as O

pytestmark = pytest.mark.gpu
gu = pytest.importorskip("gpu_util")


def _pair(shape, seed, disp=None):
    from openpiv_gpu import synthetic
    disp = disp or synthetic.sine_displacement(3.0, 2.0, 160.0)
    return synthetic.make_pair(shape, disp, synthetic.ParticleModel(), seed)


@pytest.mark.parametrize("iters,dt,scaling", [((1,), 0.5, 2.0), ((1, 1), 2.0, 0.25), ((1, 1, 1), 0.8, 3.0)])
def test_fp32_scaling_matches_oracle(iters, dt, scaling):
    from openpiv_gpu.gpu_process import piv_gpu
    shape = (320, 384)
    fa, fb = _pair(shape, 51)
    piv = piv_gpu(shape, 16 if len(iters) > 1 else 32, search_size_iters=iters, overlap_ratio=0.5, median_tol=2,
                  dt=dt, scaling_par=scaling, dtype_f="float32")
    u, v = piv(fa, fb)
    s = O.OracleSettings(min_search_size=16 if len(iters) > 1 else 32, search_size_iters=iters, dt=dt,
                         scaling_par=scaling)
    r = O.piv(fa, fb, s, keep_passes=True)
    excl = gu.pass_exclusions(fa, fb, s, r["passes"])
    rep = gu.e2e_compare(u, v, piv.val_locations, r, excl, s, tol=1e-3 * scaling / dt, final_flags=piv.outliers)
    assert rep["compared"] > 0.95 * u.size, rep
    x, y = piv.coords
    assert np.array_equal(x, r["x"]) and np.array_equal(y, r["y"])


@pytest.mark.parametrize("shape,overlap", [((384, 512), 0.5), ((320, 448), 0.75)])
def test_fp32_repeated_sizes_match_oracle(shape, overlap):
    from openpiv_gpu.gpu_process import PIVGPU
    fa, fb = _pair(shape, 52)
    kw = dict(search_size_iters=(1, 2, 2), overlap_ratio=overlap, median_tol=2.0)
    piv = PIVGPU(shape, 16, dtype_f="float32", **kw)
    assert [f.search_size for f in piv.piv_fields] == [64, 32, 32, 16, 16]
    u, v = piv.run_host(fa, fb)
    s = O.OracleSettings(min_search_size=16, **kw)
    r = O.piv(fa, fb, s, keep_passes=True)
    assert [st.geom.search_size for st in r["passes"]] == [64, 32, 32, 16, 16]
    excl = gu.pass_exclusions(fa, fb, s, r["passes"])
    rep = gu.e2e_compare(u, v, piv.init_val_locations, r, excl, s, final_flags=piv.val_locations)
    assert rep["compared"] > 0.95 * u.size, rep
    for k, st in enumerate(r["passes"]):
        assert piv.outlier_counts(k)[0] == int(st.val.sum()), k
        if k < 4:
            assert abs(piv.smoothing_log10_s(k) - st.smooth_p) < 5e-3, (k, piv.smoothing_log10_s(k), st.smooth_p)


def test_rejected_mask_leaves_the_plan_as_it_was():
    """A mask the plan cannot take (weighted smoothing of a field wider than
    the supported maximum) raises before anything changes: the plan keeps
    running unmasked, bit for bit as before."""
    from openpiv_gpu import _native
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (48, 8208)   # 16 px windows at spacing 8: a 5 x 1025 smoothed field
    fa, fb = _pair(shape, 53, None)
    piv = PIVGPU(shape, 16, search_size_iters=(2,), overlap_ratio=0.5, median_tol=2.0, dtype_f="float32")
    assert piv.piv_fields[0].field_shape[1] > 1024
    u0, v0 = piv.run_host(fa, fb)
    bad = np.zeros(shape, bool)
    bad[20:30, 100:140] = True
    with pytest.raises(_native.NativeError):
        piv._plan.set_mask(bad)
    u1, v1 = piv.run_host(fa, fb)
    assert np.array_equal(u0, u1) and np.array_equal(v0, v1)
    # a class built with that mask raises, and keeps no half-made plan
    masked = PIVGPU(shape, 16, search_size_iters=(2,), overlap_ratio=0.5, median_tol=2.0, dtype_f="float32",
                    mask=bad)
    for _ in range(2):
        with pytest.raises(_native.NativeError):
            masked.run_host(fa, fb)
        assert masked._plan is None
