"""Pins the CPU oracle (oracle/piv_oracle.py) to facts that do not depend on
the oracle's own formulas: closed-form geometry values, brute-force loops,
known answers, textbook library routines and invariants.  The reference ships
no golden vectors, so these checks are what the oracle is trusted on."""

import numpy as np
import pytest
import scipy.fft as sfft
from scipy.interpolate import RegularGridInterpolator
from scipy.sparse.linalg import spsolve

from oracle import piv_oracle as O


# --------------------------------------------------------------------------
# geometry: values of the window grid derived by hand for known workloads
# --------------------------------------------------------------------------
@pytest.mark.parametrize("shape,ss,ratio,grid,offset,x0,x_last", [
    ((512, 512), 32, 0.5, (31, 31), (0, 0), 16.0, 496.0),
    ((2048, 2048), 64, 0.5, (63, 63), (0, 0), 32.0, 2016.0),
    ((2048, 2048), 32, 0.5, (127, 127), (0, 0), 16.0, 2032.0),
    ((2048, 2048), 16, 0.5, (255, 255), (0, 0), 8.0, 2040.0),
    ((512, 1608), 64, 0.5, (15, 49), (4, 0), 36.0, 36.0 + 48 * 32),
    ((512, 1608), 32, 0.5, (31, 99), (4, 0), 20.0, 20.0 + 98 * 16),
    ((512, 1608), 16, 0.5, (63, 200), (0, 0), 8.0, 8.0 + 199 * 8),
    ((2048, 2048), 48, 0.75, (167, 167), (4, 4), 28.0, 28.0 + 166 * 12),
    ((2048, 2048), 24, 0.75, (338, 338), (1, 1), 13.0, 13.0 + 337 * 6),
])
def test_geometry_known_values(shape, ss, ratio, grid, offset, x0, x_last):
    g = O.field_geometry(shape, ss, ratio)
    assert g.field_shape == grid
    assert g.offset == offset
    assert g.x.dtype == np.float32
    assert g.x[0, 0] == x0 and g.x[0, -1] == x_last
    assert np.all(g.y[:, 0] == g.y_grid)


def test_pass_schedule_and_fft_size():
    assert O.pass_sizes(16, (1, 1, 1)) == [64, 32, 16]
    assert O.pass_sizes(16, (1, 2, 2)) == [64, 32, 32, 16, 16]
    assert O.pass_sizes(32, 1) == [32]
    assert [O.fft_size(s) for s in (16, 24, 32, 48, 64)] == [32, 64, 64, 128, 128]
    assert O.fft_size(32, 1) == 32


def test_get_stack_brute_force():
    rng = np.random.default_rng(0)
    f = rng.integers(0, 255, (40, 53)).astype(np.float32)
    g = O.field_geometry(f.shape, 8, 0.5)
    st = O.get_stack(f, g)
    m, n = g.field_shape
    x_off, y_off = g.offset
    for k in range(m * n):
        for i in range(8):
            for j in range(8):
                assert st[k, i, j] == f[(k // n) * g.spacing + y_off + i, (k % n) * g.spacing + x_off + j]


# --------------------------------------------------------------------------
# correlation
# --------------------------------------------------------------------------
def _direct_planes(wa, wb, F):
    """O(F^4) circular cross-correlation of mean-removed, zero-padded windows,
    indexed like the fftshift-ed plane: c[i, j] at lag (i - F/2, j - F/2)."""
    n, ss, _ = wa.shape
    pad = (F - ss) // 2
    out = np.zeros((n, F, F))
    for k in range(n):
        a = np.zeros((F, F))
        b = np.zeros((F, F))
        a[pad:pad + ss, pad:pad + ss] = wa[k] - wa[k].mean()
        b[pad:pad + ss, pad:pad + ss] = wb[k] - wb[k].mean()
        for i in range(F):
            for j in range(F):
                dy, dx = i - F // 2, j - F // 2
                out[k, i, j] = np.sum(a * np.roll(np.roll(b, -dy, 0), -dx, 1))
    return out


def test_correlation_equals_direct_sum():
    rng = np.random.default_rng(1)
    wa = rng.random((3, 8, 8))
    wb = rng.random((3, 8, 8))
    planes = O.correlate_windows(wa, wb, 16)
    np.testing.assert_allclose(planes, _direct_planes(wa, wb, 16), rtol=0, atol=1e-10)


@pytest.mark.parametrize("dy,dx", [(3, -2), (-5, 4), (0, 7), (-7, -7)])
def test_integer_shift_known_answer(dy, dx):
    rng = np.random.default_rng(2)
    big = rng.random((64, 64))
    ss, F = 16, 32
    a = big[20:20 + ss, 20:20 + ss][None]
    b = big[20 - dy:20 - dy + ss, 20 - dx:20 - dx + ss][None]  # content moved by +d
    planes = O.correlate_windows(a.astype(np.float32), b.astype(np.float32), F)
    ip, jp = O.first_peak(planes)
    assert (ip[0] - F // 2, jp[0] - F // 2) == (dy, dx)


def test_first_peak_first_occurrence():
    c = np.zeros((2, 4, 4), np.float32)
    c[0, 1, 2] = c[0, 2, 1] = c[0, 3, 3] = 5.0
    c[1, 3, 0] = c[1, 3, 1] = 1.0
    ip, jp = O.first_peak(c)
    assert (ip[0], jp[0]) == (1, 2)
    assert (ip[1], jp[1]) == (3, 0)


def test_gaussian_subpixel_exact_on_gaussians():
    F = 32
    y0, x0, s = 14.3, 17.8, 1.4
    ii, jj = np.meshgrid(np.arange(F), np.arange(F), indexing="ij")
    plane = np.exp(-((ii - y0) ** 2 + (jj - x0) ** 2) / (2 * s * s))[None]
    ip, jp = O.first_peak(plane)
    isp, jsp = O.subpixel_peak(plane, ip, jp, "gaussian")
    assert abs(isp[0] - y0) < 1e-6 and abs(jsp[0] - x0) < 1e-6
    # border peak and non-positive neighbours keep the integer peak
    p2 = np.zeros((1, 8, 8))
    p2[0, 0, 3] = 1.0
    assert O.subpixel_peak(p2, *O.first_peak(p2))[1][0] == 3.0
    p3 = np.zeros((1, 8, 8))
    p3[0, 4, 4] = 1.0
    p3[0, 3, 4] = 0.5
    isp, jsp = O.subpixel_peak(p3, *O.first_peak(p3), method="gaussian")
    assert (isp[0], jsp[0]) == (4.0, 4.0)


# --------------------------------------------------------------------------
# interpolation and deformation
# --------------------------------------------------------------------------
def test_rgi_matches_scipy_with_extrapolation():
    rng = np.random.default_rng(3)
    yg = np.arange(5) * 16 + 8.0
    xg = np.arange(7) * 16 + 8.0
    vals = rng.standard_normal((5, 7))
    yq = rng.uniform(-10, 90, 500)
    xq = rng.uniform(-10, 120, 500)
    yq[:5] = yg[:5]
    xq[:5] = xg[2]
    ref = RegularGridInterpolator((yg, xg), vals, bounds_error=False, fill_value=None)(np.stack([yq, xq], -1))
    np.testing.assert_allclose(O.rgi_linear(yg, xg, vals, yq, xq), ref, rtol=0, atol=1e-12)


def test_warp_reproduces_affine_intensity():
    ht, wd = 40, 50
    yy, xx = np.meshgrid(np.arange(ht, dtype=np.float64), np.arange(wd, dtype=np.float64), indexing="ij")
    f = 3.0 + 0.7 * xx - 0.4 * yy
    rng = np.random.default_rng(4)
    up = rng.uniform(-3, 3, (ht, wd))
    vp = rng.uniform(-3, 3, (ht, wd))
    out = O.warp_frame(f, up, vp, 0.5, dtype=np.float64)
    expect = 3.0 + 0.7 * (xx + 0.5 * up) - 0.4 * (yy + 0.5 * vp)
    np.testing.assert_allclose(out, expect, rtol=0, atol=1e-9)


def test_warp_brute_force_float32():
    rng = np.random.default_rng(5)
    f = rng.integers(0, 4096, (12, 15)).astype(np.float32)
    up = rng.uniform(-4, 4, f.shape).astype(np.float32)
    vp = rng.uniform(-4, 4, f.shape).astype(np.float32)
    out = O.warp_frame(f, up, vp, -0.5)
    for i in range(12):
        for j in range(15):
            xi = np.float32(j + np.float64(up[i, j]) * -0.5)
            yi = np.float32(i + np.float64(vp[i, j]) * -0.5)
            xl = min(max(int(np.floor(xi)), 0), 13)
            yu = min(max(int(np.floor(yi)), 0), 10)
            ty, tx = yi - yu, xi - xl
            val = ((1 - ty) * (1 - tx) * f[yu, xl] + (1 - ty) * tx * f[yu, xl + 1]
                   + ty * (1 - tx) * f[yu + 1, xl] + ty * tx * f[yu + 1, xl + 1])
            assert abs(out[i, j] - val) <= 2e-3


# --------------------------------------------------------------------------
# validation and replacement
# --------------------------------------------------------------------------
def test_median_matches_brute_force():
    rng = np.random.default_rng(6)
    u = rng.standard_normal((6, 7))
    med = O.neighbour_median(u)
    for i in range(6):
        for j in range(7):
            vals = [u[a, b] for a in range(i - 1, i + 2) for b in range(j - 1, j + 2)
                    if (a, b) != (i, j) and 0 <= a < 6 and 0 <= b < 7]
            assert med[i, j] == pytest.approx(np.median(vals), abs=1e-12)
    v = np.zeros_like(u)
    u2 = 0.1 * u
    u2[3, 3] = 50.0
    flags = O.median_validation(u2, v, 2.0)
    assert flags[3, 3] and flags.sum() == 1


def test_spring_single_and_pair():
    rng = np.random.default_rng(7)
    u = rng.standard_normal((5, 6))
    v = rng.standard_normal((5, 6))
    val = np.zeros((5, 6), bool)
    val[2, 3] = True                       # interior: mean of the 4 neighbours
    ur, vr = O.spring_replacement(u, v, val)
    assert ur[2, 3] == pytest.approx((u[1, 3] + u[3, 3] + u[2, 2] + u[2, 4]) / 4, abs=1e-12)
    val = np.zeros((5, 6), bool)
    val[0, 0] = True                       # corner: mean of 2 neighbours
    ur, _ = O.spring_replacement(u, v, val)
    assert ur[0, 0] == pytest.approx((u[0, 1] + u[1, 0]) / 2, abs=1e-12)
    val = np.zeros((5, 6), bool)
    val[2, 2] = val[2, 3] = True           # two linked interior outliers
    s1 = u[1, 2] + u[3, 2] + u[2, 1]
    s2 = u[1, 3] + u[3, 3] + u[2, 4]
    ur, _ = O.spring_replacement(u, v, val)
    assert ur[2, 2] == pytest.approx((4 * s1 + s2) / 15, abs=1e-12)
    assert ur[2, 3] == pytest.approx((4 * s2 + s1) / 15, abs=1e-12)


def test_spring_matrix_solve_equals_laplace_fill():
    rng = np.random.default_rng(8)
    u = rng.standard_normal((12, 11))
    val = rng.random((12, 11)) < 0.25
    link, rhs, (iv, jv) = O.spring_system(val, [u])
    x = spsolve(link.tocsc(), rhs[0])
    f = u.copy()
    f[iv, jv] = x
    # every replaced value is the mean of its in-bounds 4-neighbours
    for a, b in zip(iv, jv):
        nb = [f[p, q] for p, q in ((a - 1, b), (a + 1, b), (a, b - 1), (a, b + 1)) if 0 <= p < 12 and 0 <= q < 11]
        assert f[a, b] == pytest.approx(np.mean(nb), abs=1e-10)


# --------------------------------------------------------------------------
# smoothn
# --------------------------------------------------------------------------
def _dct_matrix(n):
    k = np.arange(n)[:, None]
    i = np.arange(n)[None, :]
    c = np.cos(np.pi * (2 * i + 1) * k / (2 * n)) * np.sqrt(2.0 / n)
    c[0] /= np.sqrt(2.0)
    return c


def test_dct_is_textbook_orthonormal_dct2():
    rng = np.random.default_rng(9)
    z = rng.standard_normal((7, 9))
    cm, cn = _dct_matrix(7), _dct_matrix(9)
    np.testing.assert_allclose(O.dct2(z), cm @ z @ cn.T, atol=1e-12)
    np.testing.assert_allclose(O.idct2(O.dct2(z)), z, atol=1e-12)


def _neumann_second_difference(n):
    """1-D second-difference operator built from its stencil: row i sums
    (z_i - z_{i-1}) + (z_i - z_{i+1}) over the neighbours that exist, i.e.
    reflecting (Neumann) ends -- the penalty Garcia's smoother uses."""
    d = np.zeros((n, n))
    for i in range(n - 1):
        d[i, i] += 1.0
        d[i + 1, i + 1] += 1.0
        d[i, i + 1] -= 1.0
        d[i + 1, i] -= 1.0
    return d


def _penalty_operator(shape):
    """A = L^2 with L = D_m (x) I + I (x) D_n, the 2-D Neumann Laplacian
    acting on the row-major flattening; smoothing order 2 squares it."""
    m, n = shape
    lap = np.kron(_neumann_second_difference(m), np.eye(n)) + np.kron(np.eye(m), _neumann_second_difference(n))
    return lap @ lap


@pytest.mark.parametrize("shape", [(7, 9), (1, 12), (12, 5)])
def test_lambda_diagonalises_the_stencil_penalty(shape):
    """C^T diag(Lambda) C == (D_m (x) I + I (x) D_n)^2 for the orthonormal
    2-D DCT C (gpu_smoothn.py:165, 340-351): the eigenvalues are pinned by
    an operator built from the finite-difference stencil alone."""
    m, n = shape
    N = m * n
    cm, cn = _dct_matrix(m), _dct_matrix(n)
    c = np.kron(cm, cn)
    lam = O.smoothn_lambda(shape).reshape(-1)
    np.testing.assert_allclose(c.T @ np.diag(lam) @ c, _penalty_operator(shape), atol=1e-11)
    # and the oracle's DCT pair applies exactly that operator
    cols = [O.idct2(O.smoothn_lambda(shape) * O.dct2(e)).reshape(-1) for e in np.eye(N).reshape(N, m, n)]
    np.testing.assert_allclose(np.array(cols).T, _penalty_operator(shape), atol=1e-11)


def test_smoothn_bounds_and_fixed_s():
    p_min, p_max = O.smoothn_p_bounds((63, 63))
    assert p_min == pytest.approx(-4.078650143894, abs=1e-9)
    assert p_max == pytest.approx(4.195550285184, abs=1e-9)
    rng = np.random.default_rng(10)
    u = rng.standard_normal((9, 8))
    s = 3.7
    z = O.smoothn([u], s=s)[0]
    # penalised least squares: z minimises |z - u|^2 + s |L z|^2, so
    # (I + s L^2) z = u, solved densely with the stencil operator
    expect = np.linalg.solve(np.eye(u.size) + s * _penalty_operator(u.shape), u.reshape(-1)).reshape(u.shape)
    np.testing.assert_allclose(z, expect, atol=1e-12)
    c = np.full((9, 8), 2.5)
    np.testing.assert_allclose(O.smoothn([c], s=100.0)[0], c, atol=1e-12)


def _dense_gcv(p, fields):
    """Generalised cross-validation from the dense hat matrix
    H = (I + 10^p A)^-1: (|(I - H) y|^2 / N) / (1 - tr H / N)^2, summed over
    the fields and divided by their number (gpu_smoothn.py:353-372)."""
    shape = fields[0].shape
    N = fields[0].size
    h = np.linalg.inv(np.eye(N) + 10.0 ** p * _penalty_operator(shape))
    rss = sum(np.sum(((np.eye(N) - h) @ y.reshape(-1)) ** 2) for y in fields)
    return rss / N / (1.0 - np.trace(h) / N) ** 2 / len(fields)


@pytest.mark.parametrize("shape", [(7, 9), (10, 10)])
def test_gcv_score_equals_dense_hat_matrix_gcv(shape):
    rng = np.random.default_rng(12)
    fields = [rng.standard_normal(shape), rng.standard_normal(shape) + 0.5]
    fd = [O.dct2(a) for a in fields]
    lam = O.smoothn_lambda(shape)
    for p in (-3.5, -1.0, 0.0, 0.7, 2.5, 4.0):
        assert O.gcv_score(p, fd, lam, 2) == pytest.approx(_dense_gcv(p, fields), rel=1e-9)


def test_weighted_gcv_equals_dense_expression():
    """Data-misfit branch (w_mean <= 0.9): the residual is taken at finite
    points against H (W (f - z) + z), the trace term is still tr H / N,
    and the normalisation is the finite count (gpu_smoothn.py:360-371)."""
    rng = np.random.default_rng(13)
    shape = (8, 7)
    N = 56
    f = [rng.standard_normal(shape), rng.standard_normal(shape)]
    z = [rng.standard_normal(shape), rng.standard_normal(shape)]
    w = (rng.random(shape) < 0.6).astype(float)
    finite = rng.random(shape) < 0.9
    w[~finite] = 0.0
    nf = int(finite.sum())
    fd = [O.dct2(w * (a - b) + b) for a, b in zip(f, z)]
    lam = O.smoothn_lambda(shape)
    for p in (-2.0, 0.4, 1.9):
        h = np.linalg.inv(np.eye(N) + 10.0 ** p * _penalty_operator(shape))
        rss = 0.0
        for a, b in zip(f, z):
            fh = (h @ (w * (a - b) + b).reshape(-1)).reshape(shape)
            rss += np.sum(w[finite] * (a[finite] - fh[finite]) ** 2)
        want = rss / nf / (1.0 - np.trace(h) / N) ** 2 / 2
        got = O.gcv_score_weighted(p, fd, lam, 2, f, w, finite, float(w.mean()), nf)
        assert got == pytest.approx(want, rel=1e-9)


def test_gcv_exact_minimiser_minimises_the_dense_gcv():
    from scipy.optimize import minimize_scalar
    rng = np.random.default_rng(14)
    yy, xx = np.mgrid[0:11, 0:12]
    fields = [np.sin(xx / 3.0) + 0.2 * rng.standard_normal((11, 12)),
              np.cos(yy / 2.5) + 0.2 * rng.standard_normal((11, 12))]
    p_min, p_max = O.smoothn_p_bounds((11, 12))
    grid = np.linspace(p_min, p_max, 301)
    g = [_dense_gcv(p, fields) for p in grid]
    k = int(np.argmin(g))
    r = minimize_scalar(lambda p: _dense_gcv(p, fields), bounds=(grid[max(k - 1, 0)], grid[min(k + 1, 300)]),
                        method="bounded", options={"xatol": 1e-10})
    assert abs(O.gcv_exact_minimiser(fields) - r.x) < 1e-6
    _, p = O.smoothn(fields, return_p=True)
    assert abs(p - r.x) < 1e-2


def test_gcv_choice_is_near_the_dense_minimum():
    rng = np.random.default_rng(11)
    yy, xx = np.meshgrid(np.arange(40), np.arange(40), indexing="ij")
    u = np.sin(xx / 6.0) + 0.1 * rng.standard_normal((40, 40))
    v = np.cos(yy / 5.0) + 0.1 * rng.standard_normal((40, 40))
    _, p = O.smoothn([u, v], return_p=True)
    p_exact = O.gcv_exact_minimiser([u, v])
    assert abs(p - p_exact) < 1e-2


# --------------------------------------------------------------------------
# end to end
# --------------------------------------------------------------------------
def test_end_to_end_uniform_shift_small():
    from openpiv_gpu import synthetic
    a, b, _ = synthetic.cfg_512_uint8(1)
    r = O.piv(a, b, O.OracleSettings(min_search_size=32))
    assert r["u"].shape == (31, 31) and r["u"].dtype == np.float32
    assert np.sqrt(np.mean((r["u"] - 2.35) ** 2)) < 0.1
    assert np.sqrt(np.mean((r["v"] + 1.60) ** 2)) < 0.1
