"""Helpers for GPU tests: device buffers (torch as the allocator) and thin
wrappers that call the C ABI stage entry points."""

import ctypes

import numpy as np

from openpiv_gpu import _native
from oracle import piv_oracle as O


def torch_mod():
    import torch
    return torch


def dev(a):
    torch = torch_mod()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def empty(shape, dtype):
    torch = torch_mod()
    return torch.empty(shape, dtype=dtype, device="cuda")


def vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def host(t):
    torch = torch_mod()
    torch.cuda.synchronize()
    return t.cpu().numpy()


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
    return p


def stage_correlate(fa, fb, geom, fsize, up=None, vpred=None, planes=False, subpixel=0):
    lib = _native.load_library()
    torch = torch_mod()
    m, n = geom.field_shape
    da, db = dev(fa), dev(fb)
    u = empty((m, n), torch.float32)
    v = empty((m, n), torch.float32)
    ip = empty((m, n), torch.int32)
    jp = empty((m, n), torch.int32)
    pl = empty((m * n, fsize, fsize), torch.float32) if planes else None
    dup = dev(np.asarray(up, np.float32)) if up is not None else None
    dvp = dev(np.asarray(vpred, np.float32)) if vpred is not None else None
    desc = pass_desc(geom, fsize, subpixel=subpixel)
    torch.cuda.synchronize()
    _native.check(lib.piv_stage_correlate(vp(da), vp(db), _native.in_dtype_code(fa.dtype), fa.shape[0], fa.shape[1],
                                          ctypes.byref(desc), vp(dup), vp(dvp), vp(u), vp(v), vp(ip), vp(jp), vp(pl),
                                          None))
    out = {"u": host(u), "v": host(v), "i_peak": host(ip), "j_peak": host(jp)}
    if planes:
        out["corr"] = host(pl)
    return out


def stage_deform(fa, fb, geom, up, vpred, deforming_par=0.5):
    lib = _native.load_library()
    torch = torch_mod()
    da, db = dev(fa), dev(fb)
    oa = empty(fa.shape, torch.float32)
    ob = empty(fa.shape, torch.float32)
    desc = pass_desc(geom, O.fft_size(geom.search_size), deforming_par=deforming_par)
    dup, dvp = dev(np.asarray(up, np.float32)), dev(np.asarray(vpred, np.float32))
    torch.cuda.synchronize()
    _native.check(lib.piv_stage_deform(vp(da), vp(db), _native.in_dtype_code(fa.dtype), fa.shape[0], fa.shape[1],
                                       ctypes.byref(desc), vp(dup), vp(dvp), vp(oa), vp(ob), None))
    return host(oa), host(ob)


def stage_grid_interp(g_from, g_to, u, v):
    lib = _native.load_library()
    torch = torch_mod()
    d_from = pass_desc(g_from, O.fft_size(g_from.search_size))
    d_to = pass_desc(g_to, O.fft_size(g_to.search_size))
    du, dv = dev(np.asarray(u, np.float32)), dev(np.asarray(v, np.float32))
    uo = empty(g_to.field_shape, torch.float32)
    vo = empty(g_to.field_shape, torch.float32)
    torch.cuda.synchronize()
    _native.check(lib.piv_stage_grid_interp(ctypes.byref(d_from), ctypes.byref(d_to), vp(du), vp(dv), vp(uo), vp(vo),
                                            None))
    return host(uo), host(vo)


def stage_validate(u, v, tol=2.0):
    lib = _native.load_library()
    torch = torch_mod()
    m, n = u.shape
    du, dv = dev(np.asarray(u, np.float32)), dev(np.asarray(v, np.float32))
    flags = empty((m, n), torch.uint8)
    count = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    _native.check(lib.piv_stage_median_validate(vp(du), vp(dv), m, n, float(tol), vp(flags), vp(count), None))
    return host(flags).astype(bool), int(host(count)[0])


def stage_replace(u, v, flags):
    lib = _native.load_library()
    torch = torch_mod()
    m, n = u.shape
    du, dv = dev(np.asarray(u, np.float32)), dev(np.asarray(v, np.float32))
    df = dev(np.asarray(flags, np.uint8))
    count = dev(np.array([int(np.count_nonzero(flags))], np.int32))
    torch.cuda.synchronize()
    _native.check(lib.piv_stage_spring_replace(vp(du), vp(dv), vp(df), vp(count), m, n, None))
    return host(du), host(dv)


def stage_smoothn(u, v, s=None):
    lib = _native.load_library()
    torch = torch_mod()
    m, n = u.shape
    du, dv = dev(np.asarray(u, np.float32)), dev(np.asarray(v, np.float32))
    p = torch.zeros(1, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _native.check(lib.piv_stage_smoothn(vp(du), vp(dv), m, n, float(s) if s else 0.0, vp(p), None))
    return host(du), host(dv), float(host(p)[0])


def near_tie_mask(planes, rel=1e-5):
    """Windows whose two largest correlation values are within ``rel`` of
    each other (integer peak not reproducible across FFT implementations)."""
    flat = planes.reshape(planes.shape[0], -1)
    part = np.partition(flat, -2, axis=1)
    top, second = part[:, -1], part[:, -2]
    return (top - second) <= rel * np.maximum(np.abs(top), 1e-30)
