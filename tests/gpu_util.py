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
               revalidate=False, s2n_method="peak2peak", s2n_size=2):
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
