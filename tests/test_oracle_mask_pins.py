"""Pins of the oracle's masked processing: the field mask, masked frames
and deformation (gpu_process.py:682-687, 949-951, 1461-1506), and the
weighted smoother (gpu_smoothn.py:123-297, 353-372) -- nearest-finite fill
against brute force, the relaxed iteration against a dense solve of the
penalised least-squares normal equations, the data-misfit GCV branch
against the spectral one (Parseval), and the coarse start's coefficient
truncation."""

import math

import numpy as np
import pytest

from oracle import piv_oracle as O


def _brute_nearest(missing):
    """Nearest non-missing element by exact squared distance; equal
    distances go to the smaller column, then the smaller row."""
    m, n = missing.shape
    fi, fj = np.nonzero(~missing)
    key = fj.astype(np.int64) * (m + 1) + fi
    out = np.stack(np.mgrid[0:m, 0:n]).astype(np.int64)
    for i, j in np.argwhere(missing):
        d = (fi - i) ** 2 + (fj - j) ** 2
        c = np.nonzero(d == d.min())[0]
        k = c[np.argmin(key[c])]
        out[:, i, j] = (fi[k], fj[k])
    return out


@pytest.mark.parametrize("kind", ["random", "disk", "rect"])
def test_nearest_finite_index_matches_brute_force(kind):
    rng = np.random.default_rng({"random": 1, "disk": 2, "rect": 3}[kind])
    for _ in range(40):
        m, n = (int(x) for x in rng.integers(4, 40, 2))
        yy, xx = np.mgrid[0:m, 0:n]
        if kind == "random":
            missing = rng.random((m, n)) < rng.uniform(0.2, 0.9)
        elif kind == "disk":
            missing = (yy - rng.uniform(0, m)) ** 2 + (xx - rng.uniform(0, n)) ** 2 < rng.uniform(1, 15) ** 2
        else:
            missing = np.zeros((m, n), bool)
            a, b = sorted(rng.integers(0, m, 2))
            c, d = sorted(rng.integers(0, n, 2))
            missing[a:b + 1, c:d + 1] = True
        if missing.all() or not missing.any():
            continue
        assert np.array_equal(O.nearest_finite_index(missing), _brute_nearest(missing))


def _pls_matrix(shape):
    """A = L^2, L the 2-D Laplacian with reflecting ends built from its
    finite-difference stencil (no DCT, no eigenvalue formula)."""
    m, n = shape

    def d1(k):
        d = np.zeros((k, k))
        for i in range(k - 1):
            d[i, i] += 1.0
            d[i + 1, i + 1] += 1.0
            d[i, i + 1] -= 1.0
            d[i + 1, i] -= 1.0
        return d

    lap = np.kron(d1(m), np.eye(n)) + np.kron(np.eye(m), d1(n))
    return lap @ lap


def test_weighted_iteration_converges_to_dense_pls_solution():
    """With a fixed s the relaxed iteration's fixed point solves
    (W + s A) z = W f; masked entries come back 0."""
    rng = np.random.default_rng(4)
    m, n = 9, 11
    yy, xx = np.mgrid[0:m, 0:n]
    u = np.sin(xx / 3.0) + 0.2 * rng.standard_normal((m, n))
    v = np.cos(yy / 4.0) + 0.2 * rng.standard_normal((m, n))
    mask = (yy - 4) ** 2 + (xx - 5) ** 2 < 5
    u[mask] = np.nan
    v[mask] = np.nan
    u[0, 0] = np.inf                     # a missing value outside the mask: weight 0, kept smoothed
    s = 2.5
    z, info = O.smoothn([u, v], s=s, mask=mask, tol_z=1e-13, max_iters=20000, return_info=True)
    assert info["weighted"] and info["iterations"] < 20000
    W = np.diag((np.isfinite(u) & np.isfinite(v) & ~mask).reshape(-1).astype(float))
    A = _pls_matrix((m, n))
    for f, zz in zip((u, v), z):
        ff = np.where(np.isfinite(f), f, 0.0).reshape(-1)
        ref = np.linalg.solve(W + s * A, W @ ff).reshape(m, n)
        ref[mask] = 0.0
        assert np.allclose(zz, ref, atol=1e-8)


def test_unweighted_path_is_one_closed_form_pass():
    rng = np.random.default_rng(5)
    f = [rng.standard_normal((8, 10)), rng.standard_normal((8, 10))]
    z, info = O.smoothn(f, s=3.0, return_info=True)
    assert info["iterations"] == 1 and not info["weighted"]
    gamma = 1.0 / (1.0 + 3.0 * O.smoothn_lambda((8, 10)))
    assert np.allclose(z[0], O.idct2(gamma * O.dct2(f[0])))


def test_slow_gcv_branch_equals_spectral_branch_for_unit_weights():
    """sum w (f - IDCT(Gamma F))^2 = ||F (Gamma - 1)||^2 when w = 1 and
    F = DCT(f) (Parseval for the orthonormal DCT)."""
    rng = np.random.default_rng(6)
    shape = (12, 7)
    f = [rng.standard_normal(shape), rng.standard_normal(shape)]
    fd = [O.dct2(a) for a in f]
    lam = O.smoothn_lambda(shape)
    w = np.ones(shape)
    fin = np.ones(shape, bool)
    for p in (-2.0, 0.3, 1.7):
        slow = O.gcv_score_weighted(p, fd, lam, 2, f, w, fin, 0.5, fin.sum())
        fast = O.gcv_score_weighted(p, fd, lam, 2, f, w, fin, 1.0, fin.sum())
        assert math.isclose(slow, fast, rel_tol=1e-10)
        assert math.isclose(fast, O.gcv_score(p, fd, lam, 2), rel_tol=1e-12)


def test_coarse_start_truncates_the_joint_high_block_only():
    rng = np.random.default_rng(7)
    f = [rng.standard_normal((23, 31))]
    z = O.smoothn_init_z(f)[0]
    d0, d1 = O.dct2(f[0]), O.dct2(z)
    n0, n1 = math.ceil(23 / 10) + 1, math.ceil(31 / 10) + 1
    assert np.allclose(d1[n0:, n1:], 0.0, atol=1e-12)
    keep = np.ones((23, 31), bool)
    keep[n0:, n1:] = False
    assert np.allclose(d1[keep], d0[keep])


def test_field_mask_frames_and_warp():
    shape = (64, 80)
    g = O.field_geometry(shape, 16, 0.5)
    rng = np.random.default_rng(8)
    mask = rng.random(shape) < 0.3
    fm = O.field_mask(g, mask)
    for i in range(g.field_shape[0]):
        for j in range(g.field_shape[1]):
            assert fm[i, j] == mask[int(g.y[i, j]), int(g.x[i, j])]
    a = rng.random(shape).astype(np.float32)
    b = rng.random(shape).astype(np.float32)
    ma, mb = O.mask_frames(a, b, mask)
    assert np.all(ma[mask] == 0) and np.array_equal(ma[~mask], a[~mask]) and np.all(mb[mask] == 0)
    up = np.full(g.field_shape, 0.7)
    vp = np.full(g.field_shape, -0.4)
    wa, wb = O.deform_frames(ma, mb, g, up, vp, 0.5, np.float32, mask)
    ua, ub = O.deform_frames(ma, mb, g, up, vp, 0.5, np.float32)
    assert np.all(wa[mask] == 0) and np.all(wb[mask] == 0)
    assert np.array_equal(wa[~mask], ua[~mask]) and np.array_equal(wb[~mask], ub[~mask])


def test_masked_pipeline_invariants():
    """Masked windows give NaN vectors and are never flagged; away from the
    mask the field follows the known displacement."""
    from openpiv_gpu import synthetic
    shape = (192, 224)
    a, b = synthetic.make_pair(shape, synthetic.uniform_displacement(2.35, -1.6), synthetic.ParticleModel(), 11)
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    mask = (yy - 96) ** 2 + (xx - 100) ** 2 < 30 ** 2
    st = O.OracleSettings(min_search_size=16, search_size_iters=(1, 1), mask=mask, s2n_tol=1.05)
    r = O.piv(a, b, st, keep_passes=True)
    g = r["passes"][-1].geom
    fm = O.field_mask(g, mask)
    assert np.array_equal(np.isnan(r["u"]), fm) and np.array_equal(np.isnan(r["v"]), fm)
    assert not r["val_locations"][fm].any() and not r["outliers"][fm].any()
    assert np.all(r["passes"][-1].corr["s2n"][fm] == 0)
    # the smoothed predictor is 0 inside the mask
    p0 = r["passes"][0]
    assert np.all(p0.u_out[O.field_mask(p0.geom, mask)] == 0)
    # vectors whose window does not reach the mask
    far = np.zeros_like(fm)
    for i in range(fm.shape[0]):
        for j in range(fm.shape[1]):
            y0, x0 = int(g.y[i, j]) - 8, int(g.x[i, j]) - 8
            far[i, j] = not mask[max(y0 - 4, 0):y0 + 20, max(x0 - 4, 0):x0 + 20].any()
    err = np.hypot(r["u"][far] - 2.35, r["v"][far] + 1.6)
    assert far.sum() > 20 and np.sqrt(np.mean(err ** 2)) < 0.15 and err.max() < 0.5
