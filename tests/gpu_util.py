"""Helpers for GPU tests: device buffers from the library's own allocator and
thin wrappers that call the C ABI stage entry points."""

import ctypes

import numpy as np

from openpiv_gpu import _native
from oracle import piv_oracle as O


def dev(a):
    return _native.DeviceBuffer.from_array(np.ascontiguousarray(a))


def empty(shape, dtype):
    return _native.DeviceBuffer(tuple(shape), dtype)


def vp(b):
    return ctypes.c_void_p(b.value) if b is not None else None


def host(b):
    return b.get()


def pass_desc(geom: O.Geometry, fsize: int, normalize=True, subpixel=0, deforming_par=0.5):
    p = _native.PassDesc()
    p.search_size = geom.search_size
    p.spacing = geom.spacing
    p.x_offset, p.y_offset = geom.offset
    p.field_m, p.field_n = geom.field_shape
    p.fft_size = fsize
    p.normalize = int(normalize or fsize != geom.search_size)
    p.subpixel = subpixel
    p.deforming_par = deforming_par
    p.validate = 1
    p.median_tol = 2.0
    p.num_replacing_iters = 1
    p.smooth = 1
    p.smoothing_par = 0.0
    nan = float("nan")
    p.s2n_method, p.s2n_size = 0, 2
    p.s2n_tol = p.mad_tol = p.mean_tol = p.rms_tol = nan
    p.replacing_method = 0
    p.revalidate = 0
    return p


def tests_desc(median_tol=2.0, s2n_tol=None, mad_tol=None, mean_tol=None, rms_tol=None, method="spring", n_iters=1,
               revalidate=False, s2n_method="peak2peak", s2n_size=2, validation_size=1, replacing_size=1):
    """A pass descriptor carrying only validation / replacement settings."""
    p = _native.PassDesc()
    nan = float("nan")
    f = lambda t: nan if t is None else float(t)
    p.validate = 1
    p.median_tol, p.s2n_tol, p.mad_tol, p.mean_tol, p.rms_tol = (f(median_tol), f(s2n_tol), f(mad_tol), f(mean_tol),
                                                                 f(rms_tol))
    p.s2n_method, p.s2n_size = _native.S2N_CODES[s2n_method], s2n_size
    p.replacing_method = _native.REPLACING_CODES[method]
    p.num_replacing_iters = n_iters
    p.revalidate = int(revalidate)
    p.validation_size = validation_size
    p.replacing_size = replacing_size
    return p


def stage_validate_all(u, v, desc, s2n=None, only=None):
    lib = _native.load_library()
    m, n = u.shape
    du, dv = dev(np.asarray(u, np.float32)), dev(np.asarray(v, np.float32))
    ds = dev(np.asarray(s2n, np.float32)) if s2n is not None else None
    do = dev(np.asarray(only, np.uint8)) if only is not None else None
    fl = empty((m, n), np.uint8)
    cnt = _native.DeviceBuffer.zeros((1,), np.int32)
    _native.check(lib.piv_stage_validate(vp(du), vp(dv), m, n, vp(ds), ctypes.byref(desc), vp(do), vp(fl), vp(cnt),
                                         None))
    return host(fl).astype(bool), int(host(cnt)[0])


def stage_replace_all(u, v, flags, desc, s2n=None):
    lib = _native.load_library()
    m, n = u.shape
    du, dv = dev(np.asarray(u, np.float32)), dev(np.asarray(v, np.float32))
    df = dev(np.asarray(flags, np.uint8))
    ds = dev(np.asarray(s2n, np.float32)) if s2n is not None else None
    _native.check(lib.piv_stage_replace(vp(du), vp(dv), vp(df), m, n, ctypes.byref(desc), vp(ds), None))
    return host(du), host(dv), host(df).astype(bool)


def stage_correlate(fa, fb, geom, fsize, up=None, vpred=None, planes=False, subpixel=0, s2n_method=None,
                    s2n_size=2, window_size=0):
    lib = _native.load_library()
    m, n = geom.field_shape
    da, db = dev(fa), dev(fb)
    u = empty((m, n), np.float32)
    v = empty((m, n), np.float32)
    ip = empty((m, n), np.int32)
    jp = empty((m, n), np.int32)
    pl = empty((m * n, fsize, fsize), np.float32) if planes else None
    sn = empty((m, n), np.float32) if s2n_method is not None else None
    dup = dev(np.asarray(up, np.float32)) if up is not None else None
    dvp = dev(np.asarray(vpred, np.float32)) if vpred is not None else None
    desc = pass_desc(geom, fsize, subpixel=subpixel)
    desc.window_size = window_size
    if s2n_method is not None:
        desc.s2n_method = _native.S2N_CODES[s2n_method]
        desc.s2n_size = s2n_size
    _native.check(lib.piv_stage_correlate(vp(da), vp(db), _native.in_dtype_code(fa.dtype), fa.shape[0], fa.shape[1],
                                          ctypes.byref(desc), vp(dup), vp(dvp), vp(u), vp(v), vp(ip), vp(jp), vp(pl),
                                          vp(sn), None))
    out = {"u": host(u), "v": host(v), "i_peak": host(ip), "j_peak": host(jp)}
    if planes:
        out["corr"] = host(pl)
    if sn is not None:
        out["s2n"] = host(sn)
    return out


def stage_correlate64(fa, fb, geom, fsize, up=None, vpred=None, subpixel=0, s2n_method=None, s2n_size=2, window_size=0):
    lib = _native.load_library()
    m, n = geom.field_shape
    da, db = dev(fa), dev(fb)
    u = empty((m, n), np.float64)
    v = empty((m, n), np.float64)
    ip = empty((m, n), np.int32)
    jp = empty((m, n), np.int32)
    pl = empty((m * n, fsize, fsize), np.float64)
    sn = empty((m, n), np.float64) if s2n_method is not None else None
    dup = dev(np.asarray(up, np.float64)) if up is not None else None
    dvp = dev(np.asarray(vpred, np.float64)) if vpred is not None else None
    desc = pass_desc(geom, fsize, subpixel=subpixel)
    desc.window_size = window_size
    if s2n_method is not None:
        desc.s2n_method = _native.S2N_CODES[s2n_method]
        desc.s2n_size = s2n_size
    _native.check(lib.piv_stage_correlate64(vp(da), vp(db), _native.in_dtype_code(fa.dtype), fa.shape[0],
                                            fa.shape[1], ctypes.byref(desc), vp(dup), vp(dvp), vp(u),
                                            vp(v), vp(ip), vp(jp), vp(pl), vp(sn), None))
    out = {"u": host(u), "v": host(v), "i_peak": host(ip), "j_peak": host(jp), "corr": host(pl)}
    if sn is not None:
        out["s2n"] = host(sn)
    return out


def stage_deform(fa, fb, geom, up, vpred, deforming_par=0.5):
    lib = _native.load_library()
    da, db = dev(fa), dev(fb)
    oa = empty(fa.shape, np.float32)
    ob = empty(fa.shape, np.float32)
    desc = pass_desc(geom, O.fft_size(geom.search_size), deforming_par=deforming_par)
    dup, dvp = dev(np.asarray(up, np.float32)), dev(np.asarray(vpred, np.float32))
    _native.check(lib.piv_stage_deform(vp(da), vp(db), _native.in_dtype_code(fa.dtype), fa.shape[0], fa.shape[1],
                                       ctypes.byref(desc), vp(dup), vp(dvp), vp(oa), vp(ob), None))
    return host(oa), host(ob)


def stage_grid_interp(g_from, g_to, u, v):
    lib = _native.load_library()
    d_from = pass_desc(g_from, O.fft_size(g_from.search_size))
    d_to = pass_desc(g_to, O.fft_size(g_to.search_size))
    du, dv = dev(np.asarray(u, np.float32)), dev(np.asarray(v, np.float32))
    uo = empty(g_to.field_shape, np.float32)
    vo = empty(g_to.field_shape, np.float32)
    _native.check(lib.piv_stage_grid_interp(ctypes.byref(d_from), ctypes.byref(d_to), vp(du), vp(dv), vp(uo), vp(vo),
                                            None))
    return host(uo), host(vo)


def stage_validate(u, v, tol=2.0):
    lib = _native.load_library()
    m, n = u.shape
    du, dv = dev(np.asarray(u, np.float32)), dev(np.asarray(v, np.float32))
    flags = empty((m, n), np.uint8)
    count = _native.DeviceBuffer.zeros((1,), np.int32)
    _native.check(lib.piv_stage_median_validate(vp(du), vp(dv), m, n, float(tol), vp(flags), vp(count), None))
    return host(flags).astype(bool), int(host(count)[0])


def stage_replace(u, v, flags):
    lib = _native.load_library()
    m, n = u.shape
    du, dv = dev(np.asarray(u, np.float32)), dev(np.asarray(v, np.float32))
    df = dev(np.asarray(flags, np.uint8))
    count = dev(np.array([int(np.count_nonzero(flags))], np.int32))
    _native.check(lib.piv_stage_spring_replace(vp(du), vp(dv), vp(df), vp(count), m, n, None))
    return host(du), host(dv)


def stage_smoothn(u, v, s=None):
    lib = _native.load_library()
    m, n = u.shape
    du, dv = dev(np.asarray(u, np.float32)), dev(np.asarray(v, np.float32))
    p = _native.DeviceBuffer.zeros((1,), np.float64)
    _native.check(lib.piv_stage_smoothn(vp(du), vp(dv), m, n, float(s) if s else 0.0, vp(p), None))
    return host(du), host(dv), float(host(p)[0])


def near_tie_mask(planes, rel=1e-5):
    """Windows whose two largest correlation values are within ``rel`` of
    each other (integer peak not reproducible across FFT implementations)."""
    flat = planes.reshape(planes.shape[0], -1)
    part = np.partition(flat, -2, axis=1)
    top, second = part[:, -1], part[:, -2]
    return (top - second) <= rel * np.maximum(np.abs(top), 1e-30)


def fit_sign_ambiguous(planes, i_peak, j_peak, rel=1e-5):
    """Windows where one of the five samples of the 3-point Gaussian fit is
    within ``rel`` (of the peak) of zero: whether the fit runs (all five > 0)
    then depends on rounding, and the result can jump by up to 0.5 px."""
    n, F, _ = planes.shape
    ip = i_peak.reshape(-1)
    jp = j_peak.reshape(-1)
    out = np.zeros(n, bool)
    for k in range(n):
        i, j = ip[k], jp[k]
        if 0 < i < F - 1 and 0 < j < F - 1:
            c = planes[k, i, j]
            vals = np.array([c, planes[k, i - 1, j], planes[k, i + 1, j], planes[k, i, j - 1], planes[k, i, j + 1]])
            out[k] = np.any(np.abs(vals) <= rel * max(abs(c), 1e-30))
    return out


def stack_geometry(n, ws):
    """Geometry that reads an (n, ws, ws) window stack laid out as an
    (n*ws, ws) frame: window k at rows k*ws .. k*ws+ws-1."""
    g = O.Geometry((n * ws, ws), ws, 0, ws, (n, 1), (0, 0), np.zeros((n, 1), np.float32),
                   np.zeros((n, 1), np.float32))
    return g


def stack_correlate_fft(wa, wb, fsize):
    n, ws, _ = wa.shape
    g = stack_geometry(n, ws)
    out = stage_correlate(wa.reshape(n * ws, ws), wb.reshape(n * ws, ws), g, fsize)
    return {k: v.reshape(-1) for k, v in out.items()}


def stack_correlate_dft(wa, wb, fsize, precision, subpixel=0, planes=False):
    lib = _native.load_library()
    n, ws, _ = wa.shape
    da, db = dev(np.ascontiguousarray(wa, np.float32)), dev(np.ascontiguousarray(wb, np.float32))
    u = empty((n,), np.float32)
    v = empty((n,), np.float32)
    ip = empty((n,), np.int32)
    jp = empty((n,), np.int32)
    pl = empty((n, fsize, fsize), np.float32) if planes else None
    _native.check(lib.piv_stage_correlate_dft(vp(da), vp(db), n, ws, fsize, precision, subpixel, vp(u), vp(v),
                                              vp(ip), vp(jp), vp(pl), None))
    out = {"u": host(u), "v": host(v), "i_peak": host(ip), "j_peak": host(jp)}
    if planes:
        out["corr"] = host(pl)
    return out


def threshold_band(u, v, s2n, settings, rel=1e-5, s2n_rel=1e-4, only=None):
    """Vectors whose validation decision is not reproducible across
    implementations: within ``rel`` (relative) of a median / MAD / mean / RMS
    threshold, or within ``s2n_rel`` of the S2N threshold (the S2N value
    itself differs by FFT rounding: ~1e-5 relative for peak2peak, a few % for
    the two mean methods).  ``settings``: an OracleSettings."""
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    near = np.zeros(u.shape, bool)

    def close(x, t):
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.abs(x - t) <= rel * np.maximum(np.abs(t), 1e-12)

    s = settings
    if s.median_tol is not None or s.mad_tol is not None or s.mean_tol is not None:
        for f in (u, v):
            med, mad, mean = O.neighbour_stats(f, getattr(s, "validation_size", 1))
            if s.median_tol is not None:
                near |= close(np.abs(f - med), s.median_tol)
            if s.mad_tol is not None:
                near |= close(np.abs(f - med), s.mad_tol * mad)
            if s.mean_tol is not None:
                near |= close(np.abs(f - mean), s.mean_tol)
    if s.rms_tol is not None:
        for f in (u, v):
            with np.errstate(invalid="ignore"):
                mu = np.nanmean(f)
                sd = np.nanstd(f)
            near |= close(np.abs(f - mu), s.rms_tol * sd)
    if s.s2n_tol is not None and s2n is not None:
        with np.errstate(invalid="ignore", divide="ignore"):
            near |= np.abs(np.asarray(s2n, np.float64).reshape(u.shape) / s.s2n_tol - 1.0) <= s2n_rel
    if only is not None:
        near &= only
    return near


def pass_exclusions(frame_a, frame_b, settings, passes, mode="faithful32", s2n_rel=1e-4):
    """The oracle's own record of which vectors of each pass an FFT
    implementation other than its own may legitimately decide differently:

    * ``tie``: windows whose two largest correlation values lie within 1e-5
      (relative) of each other, so the integer peak can flip;
    * ``fit``: windows where one of the five Gaussian-fit samples lies within
      1e-5 of zero, so whether the fit runs can flip;
    * ``band``: vectors within the validation thresholds' bands
      (threshold_band), evaluated on every field the pass validates (before
      replacement, and after it when the last pass re-validates).

    ``passes``: the oracle's PassState list of the same run (keep_passes)."""
    dtype = np.float32 if mode == "faithful32" else np.float64
    out = []
    for k, st in enumerate(passes):
        fa = np.asarray(frame_a, dtype)
        fb = np.asarray(frame_b, dtype)
        if settings.mask is not None:
            fa, fb = O.mask_frames(fa, fb, np.asarray(settings.mask, bool))
        if st.up is not None:
            fa, fb = O.deform_frames(fa, fb, st.geom, st.up, st.vp, settings.deforming_par, dtype,
                                     None if settings.mask is None else np.asarray(settings.mask, bool))
        ws = None
        if k == 0 and settings.shrink_ratio != 1:
            ws = O.shrunk_window(st.geom.search_size, settings.shrink_ratio)
        r = O.correlate_pass(fa, fb, st.geom, settings.n_fft, settings.normalize, settings.subpixel_method, dtype,
                             return_planes=True, window_size=ws)
        shape = st.geom.field_shape
        tie = near_tie_mask(r["corr"]).reshape(shape)
        fit = fit_sign_ambiguous(r["corr"], r["i_peak"], r["j_peak"]).reshape(shape)
        fm = O.field_mask(st.geom, settings.mask)
        keep = None if fm is None else ~fm
        s2n = st.corr.get("s2n")
        band = threshold_band(st.u_pre, st.v_pre, s2n, settings, s2n_rel=s2n_rel, only=keep)
        if st.outliers is not None:
            band |= threshold_band(st.u_rep, st.v_rep, s2n, settings, s2n_rel=s2n_rel, only=keep)
        if keep is not None:
            tie &= keep
            fit &= keep
        out.append({"tie": tie, "fit": fit, "band": band, "any": tie | fit | band})
    return out


def dilate(mask, steps, size=1):
    """Grow a boolean field by ``steps`` rings of width ``size`` (the reach
    of a (2 size + 1)^2 validation / replacement step)."""
    out = np.asarray(mask, bool).copy()
    w = 2 * size + 1
    for _ in range(max(0, int(steps))):
        p = np.pad(out, size)
        grown = np.zeros_like(out)
        for di in range(w):
            for dj in range(w):
                grown |= p[di:di + out.shape[0], dj:dj + out.shape[1]]
        out = grown
    return out


def _spring_clusters(flags, size):
    """Labels of the outlier clusters the spring system couples: links join
    vectors ``size`` apart in a row or column, so clusters are 4-connected
    components of each residue sublattice flags[a::size, b::size]."""
    from scipy import ndimage
    lab = np.zeros(flags.shape, np.int64)
    base = 0
    for a in range(size):
        for b in range(size):
            sub, n = ndimage.label(flags[a::size, b::size])
            lab[a::size, b::size] = np.where(sub > 0, sub + base, 0)
            base += n
    return lab


def _reach(seed, flags, settings):
    """Vectors whose replaced value can depend on the vectors in ``seed``:
    ``num_replacing_iters`` + 1 rings for the median / mean replacement
    (each iteration reads a (2 size + 1)^2 neighbourhood, revalidation one of
    validation_size), and for the spring replacement every outlier cluster
    (_spring_clusters) next to the seed."""
    iters = max(1, int(settings.num_replacing_iters))
    rs = int(getattr(settings, "replacing_size", 1))
    size = max(rs, int(getattr(settings, "validation_size", 1)))
    out = dilate(seed, iters + 1, size)
    if settings.replacing_method == "spring" and flags.any():
        lab = _spring_clusters(flags, rs)
        hit = np.unique(lab[dilate(seed, 1, rs) & flags])
        out |= np.isin(lab, hit[hit > 0])
        out = dilate(out, 1, size)
    return out


def e2e_compare(u, v, init_flags, ref, excl, settings, tol=1e-3, final_flags=None):
    """End-to-end comparison with every excused vector accounted for.

    * passes before the last must have no non-reproducible vector at all
      (``excl[k]["any"]`` empty, from pass_exclusions): the inputs are chosen
      so, and the assert says so when they are not;
    * last pass: a pre-replacement flag may differ only at a band vector or
      next to a near-tie / fit-ambiguous window (its value moves the
      neighbourhood median); u and v must agree to ``tol`` everywhere except
      within the replacement's reach of a differing flag, a near-tie or a
      fit-ambiguous window (_reach);
    * post-replacement flags (``final_flags``), when given, may differ only
      where u, v are excused or at band vectors.

    Returns a short report of what was excused."""
    for k, e in enumerate(excl[:-1]):
        assert not e["any"].any(), (f"pass {k}: {int(e['tie'].sum())} near-tie, {int(e['fit'].sum())} "
                                    f"fit-ambiguous, {int(e['band'].sum())} band vectors: pick inputs without them")
    last = excl[-1]
    ref_init = np.asarray(ref["val_locations"], bool)
    got_init = np.asarray(init_flags, bool)
    tie_fit = last["tie"] | last["fit"]
    flag_diff = got_init != ref_init
    allowed = last["band"] | dilate(tie_fit, 1, int(getattr(settings, "validation_size", 1)))
    assert not (flag_diff & ~allowed).any(), ("pre-replacement flags differ outside the bands",
                                              np.argwhere(flag_diff & ~allowed)[:10])
    excused = _reach(flag_diff | tie_fit, ref_init | got_init, settings)
    du = np.abs(np.asarray(u, np.float64) - ref["u"])
    dv = np.abs(np.asarray(v, np.float64) - ref["v"])
    fin = np.isfinite(ref["u"]) & np.isfinite(ref["v"])
    assert np.array_equal(np.isfinite(u) & np.isfinite(v), fin), "NaN pattern differs"
    ok = fin & ~excused
    err = np.maximum(du, dv)
    assert err[ok].max(initial=0.0) <= tol, ("u/v differ outside the excused vectors",
                                             float(err[ok].max()), np.argwhere(ok & (err > tol))[:10])
    if final_flags is not None and ref.get("outliers") is not None:
        fd = np.asarray(final_flags, bool) != np.asarray(ref["outliers"], bool)
        assert not (fd & ~(excused | last["band"])).any(), "post-replacement flags differ outside the excused vectors"
    return {"flag_diff": int(flag_diff.sum()), "tie": int(last["tie"].sum()), "fit": int(last["fit"].sum()),
            "band": int(last["band"].sum()), "excused": int(excused.sum()), "compared": int(ok.sum())}
