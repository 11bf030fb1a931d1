"""Pins of the oracle's S2N, full validation and replacement loop
(gpu_process.py:744-834, 1312-1365; gpu_validation.py:84-415) by hand-built
planes, brute-force neighbourhood loops and hand-stepped iterations."""

import numpy as np
import pytest

from oracle import piv_oracle as O


def _plane(F=16):
    p = np.full((1, F, F), -0.5, np.float32)
    return p


def test_s2n_peak2peak_known_answer_and_wrapped_exclusion():
    p = _plane()
    p[0, 8, 8] = 10.0
    p[0, 9, 9] = 9.0    # inside the 5x5 block around the peak: ignored
    p[0, 3, 12] = 4.0   # the second peak
    s = O.s2n_ratio(p, np.array([8]), np.array([8]), "peak2peak", 2)
    assert np.isclose(s[0], 2.5)
    # peak in the corner: the excluded block wraps to the opposite edges
    q = _plane()
    q[0, 0, 0] = 10.0
    q[0, 15, 15] = 8.0   # wrapped neighbour (-1, -1): excluded
    q[0, 14, 1] = 7.0    # wrapped (-2, +1): excluded
    q[0, 5, 5] = 2.0
    s = O.s2n_ratio(q, np.array([0]), np.array([0]), "peak2peak", 2)
    assert np.isclose(s[0], 5.0)


def test_s2n_mean_and_energy_by_hand():
    p = np.zeros((1, 4, 4), np.float64)
    p[0] = [[8, 1, 2, -1], [3, 5, 0, 0], [4, 0, -2, 0], [0, 0, 0, 0]]
    # peak2energy: mean of positive values {8, 1, 2, 3, 5, 4}
    s = O.s2n_ratio(p, np.array([0]), np.array([0]), "peak2energy")
    assert np.isclose(s[0], 8.0 / (23.0 / 6.0))
    # peak2mean: mean of values in (0, 4] = {1, 2, 3, 4}
    s = O.s2n_ratio(p, np.array([0]), np.array([0]), "peak2mean")
    assert np.isclose(s[0], 8.0 / 2.5)


def test_s2n_special_cases():
    p = np.full((3, 4, 4), -1.0)
    p[0, 0, 0] = -0.5          # first peak not positive -> 0
    p[1, 1, 1] = 2.0           # no positive second value -> inf (peak2peak: second peak -1)
    p[2, 2, 2] = 2.0           # peak2mean: nothing in (0, 1] -> NaN
    ip = np.array([0, 1, 2])
    jp = np.array([0, 1, 2])
    s = O.s2n_ratio(p, ip, jp, "peak2peak", 1)
    assert s[0] == 0 and np.isinf(s[1])
    s = O.s2n_ratio(p, ip, jp, "peak2mean")
    assert s[0] == 0 and np.isnan(s[2])


def _brute_stats(f, i, j, size=1, skip=None):
    m, n = f.shape
    nb = [f[a, b] for a in range(i - size, i + size + 1) for b in range(j - size, j + size + 1)
          if (a, b) != (i, j) and 0 <= a < m and 0 <= b < n and not (skip is not None and skip[a, b])]
    srt = sorted(nb)
    k = len(srt)
    med = srt[k // 2] if k % 2 else 0.5 * (srt[k // 2 - 1] + srt[k // 2])
    dev = sorted(abs(x - med) for x in nb)
    mad = dev[k // 2] if k % 2 else 0.5 * (dev[k // 2 - 1] + dev[k // 2])
    return med, mad, sum(nb) / k


def test_neighbour_stats_match_brute_force():
    rng = np.random.default_rng(1)
    f = rng.standard_normal((7, 9))
    med, mad, mean = O.neighbour_stats(f)
    for i in range(7):
        for j in range(9):
            bm, bd, bn = _brute_stats(f, i, j)
            assert np.isclose(med[i, j], bm) and np.isclose(mad[i, j], bd) and np.isclose(mean[i, j], bn)


def test_validate_each_test_and_only_mask():
    rng = np.random.default_rng(2)
    u = rng.standard_normal((12, 12)) * 0.1
    v = rng.standard_normal((12, 12)) * 0.1
    u[5, 5] = 3.0
    v[2, 8] = -1.0
    s2n = np.full((12, 12), 5.0)
    s2n[9, 1] = 1.1
    flags = O.validate(u, v, s2n, s2n_tol=1.3, median_tol=None)
    assert flags.sum() == 1 and flags[9, 1]
    flags = O.validate(u, v, median_tol=2.0)
    assert flags[5, 5] and not flags[2, 8]
    flags = O.validate(u, v, median_tol=None, mean_tol=0.8)
    assert flags[5, 5] and flags[2, 8]
    # RMS: global mean/std per field
    fl = O.validate(u, v, median_tol=None, rms_tol=3.0)
    ref = (np.abs(u - u.mean()) > 3 * u.std()) | (np.abs(v - v.mean()) > 3 * v.std())
    assert np.array_equal(fl, ref)
    # MAD: |f - med| > tol * mad by brute force
    fl = O.validate(u, v, median_tol=None, mad_tol=4.0)
    for i in range(12):
        for j in range(12):
            mu, du, _ = _brute_stats(u, i, j)
            mv, dv, _ = _brute_stats(v, i, j)
            assert fl[i, j] == ((abs(u[i, j] - mu) > 4 * du) or (abs(v[i, j] - mv) > 4 * dv))
    only = np.zeros((12, 12), bool)
    only[2, 8] = True
    assert O.validate(u, v, median_tol=None, mean_tol=0.8, only=only).sum() == 1


def test_neighbour_replacement_exclusion():
    u = np.arange(25, dtype=np.float64).reshape(5, 5)
    v = -u
    val = np.zeros((5, 5), bool)
    val[2, 2] = val[2, 3] = True
    ru, rv = O.neighbour_replacement(u, v, val, "mean", exclude_flagged=True)
    nb = [6, 7, 8, 11, 16, 17, 18]          # (2,2)'s neighbours without the flagged (2,3)
    assert np.isclose(ru[2, 2], np.mean(nb)) and np.isclose(rv[2, 2], -np.mean(nb))
    ru, _ = O.neighbour_replacement(u, v, val, "median", exclude_flagged=False)
    assert np.isclose(ru[2, 2], np.median([6, 7, 8, 11, 13, 16, 17, 18]))
    assert np.array_equal(ru[~val], u[~val])


def test_replacement_iterations_by_hand():
    rng = np.random.default_rng(3)
    u = rng.standard_normal((8, 8))
    v = rng.standard_normal((8, 8))
    val = np.zeros((8, 8), bool)
    val[3, 3] = val[3, 4] = val[4, 3] = True
    # two plain iterations: the second includes the replaced values
    a1 = O.neighbour_replacement(u, v, val, "median", True)
    a2 = O.neighbour_replacement(*a1, val, "median", False)
    ru, rv, rval = O.replace_outliers(u, v, val, "median", n_iters=2)
    assert np.allclose(ru, a2[0]) and np.allclose(rv, a2[1]) and np.array_equal(rval, val)
    # revalidation: the replaced vectors pass the test and are cleared
    fn = lambda a, b, only=None: O.validate(a, b, median_tol=0.5, only=only)
    u2 = np.zeros((8, 8))
    v2 = np.zeros((8, 8))
    u2[3, 3] = 9.0
    ru, rv, rval = O.replace_outliers(u2, v2, fn(u2, v2), "mean", n_iters=3, revalidate=True, validate_fn=fn)
    assert not rval.any() and ru[3, 3] == 0.0


def test_unresolved_replacements_are_reset():
    u = np.ones((6, 6))
    v = np.ones((6, 6))
    val = np.zeros((6, 6), bool)
    val[1:5, 1:5] = True                    # the inner 2x2 has only flagged neighbours
    u[1:5, 1:5] = 7.0
    ru, rv, _ = O.replace_outliers(u, v, val, "median", n_iters=1)
    assert np.all(ru[2:4, 2:4] == 7.0)      # NaN -> reset to the original
    assert np.all(ru[1, 1:5] == 1.0)        # border of the block resolved from valid neighbours


def test_shrunk_window_correlation_by_direct_sum():
    """Interrogation window (ws = 4) inside an 8 px search window, F = 16:
    A zero outside its window and centred over it, B centred over the whole
    window, then the direct circular correlation c[d] = sum_x a[x] b[x + d]."""
    rng = np.random.default_rng(9)
    wa = rng.random((2, 8, 8))
    wb = rng.random((2, 8, 8))
    got = O.correlate_windows(wa, wb, 16, True, win_pad=2)
    for k in range(2):
        a = np.zeros((16, 16))
        b = np.zeros((16, 16))
        inner = wa[k, 2:6, 2:6]
        a[6:10, 6:10] = inner - inner.mean()
        b[4:12, 4:12] = wb[k] - wb[k].mean()
        c = np.zeros((16, 16))
        for dy in range(16):
            for dx in range(16):
                c[dy, dx] = np.sum(a * np.roll(np.roll(b, -dy, 0), -dx, 1))
        assert np.allclose(got[k], np.fft.fftshift(c), atol=1e-10)
    assert O.shrunk_window(64, 0.1) == 8 and O.shrunk_window(32, 0.55) == 18 and O.shrunk_window(32, 0.6) == 20


def test_wide_neighbourhood_stats_and_replacement_by_brute_force():
    """validation_size / replacing_size 2 and 3: (2 size + 1)^2 kernels
    (gpu_validation.py:73-76, 194-205, 272-290, 393-398), with NaN entries
    and flagged neighbours left out."""
    rng = np.random.default_rng(11)
    f = rng.standard_normal((9, 11))
    f[4, 6] = np.nan
    for size in (2, 3):
        med, mad, mean = O.neighbour_stats(f, size)
        for i in range(9):
            for j in range(11):
                nb = [x for a in range(i - size, i + size + 1) for b in range(j - size, j + size + 1)
                      if (a, b) != (i, j) and 0 <= a < 9 and 0 <= b < 11 and np.isfinite(x := f[a, b])]
                srt = sorted(nb)
                k = len(srt)
                want_med = srt[k // 2] if k % 2 else 0.5 * (srt[k // 2 - 1] + srt[k // 2])
                dev = sorted(abs(x - want_med) for x in nb)
                want_mad = dev[k // 2] if k % 2 else 0.5 * (dev[k // 2 - 1] + dev[k // 2])
                assert np.isclose(med[i, j], want_med) and np.isclose(mad[i, j], want_mad)
                assert np.isclose(mean[i, j], sum(nb) / k)
        val = rng.random((9, 11)) < 0.2
        for method in ("median", "mean"):
            ru, _ = O.neighbour_replacement(f, -f, val, method, exclude_flagged=True, size=size)
            for i, j in zip(*np.nonzero(val)):
                nb = [f[a, b] for a in range(i - size, i + size + 1) for b in range(j - size, j + size + 1)
                      if (a, b) != (i, j) and 0 <= a < 9 and 0 <= b < 11 and not val[a, b] and np.isfinite(f[a, b])]
                want = np.median(nb) if method == "median" else np.mean(nb)
                assert np.isclose(ru[i, j], want), (size, method, i, j)


def test_wide_spring_splits_into_residue_sublattices():
    """Spring links size apart (gpu_validation.py:279-283, 365-371) join
    only vectors of one residue class, and the position bands of width size
    are the border rows and columns of each class: on fields at least
    2 size wide, spring replacement with size s equals size-1 replacement
    on every sublattice f[a::s, b::s]."""
    rng = np.random.default_rng(12)
    for size, shape in ((2, (13, 10)), (3, (11, 14)), (2, (4, 5))):
        u = rng.standard_normal(shape)
        v = rng.standard_normal(shape)
        val = rng.random(shape) < 0.3
        val[0, 0] = val[-1, -1] = True
        ru, rv = O.spring_replacement(u, v, val, size)
        for a in range(size):
            for b in range(size):
                su, sv = O.spring_replacement(u[a::size, b::size], v[a::size, b::size], val[a::size, b::size])
                assert np.allclose(ru[a::size, b::size], su, atol=1e-12)
                assert np.allclose(rv[a::size, b::size], sv, atol=1e-12)
    # a fully flagged class is singular and keeps its values; the others solve
    u = rng.standard_normal((8, 9))
    val = np.zeros((8, 9), bool)
    val[1::2, 0::2] = True
    val[4, 4] = True
    assert np.array_equal(O.spring_singular(val, 2), np.where(np.arange(8)[:, None] % 2 == 1, True, False)
                          & (np.arange(9)[None, :] % 2 == 0))
    ru, _ = O.spring_replacement(u, u, val, 2)
    assert np.array_equal(ru[1::2, 0::2], u[1::2, 0::2])
    assert ru[4, 4] == pytest.approx((u[2, 4] + u[6, 4] + u[4, 2] + u[4, 6]) / 4, abs=1e-12)
