"""CPU oracle for the PIV hot path: a NumPy/SciPy restatement of the
reference's ``PIVGPU`` multi-pass loop.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or the timed CPU baseline).  The product path never calls it.

Every function cites the reference file:line whose semantics it restates
(paths relative to ``openpiv_gpu/`` of the reference).  Two precision modes:

* ``"faithful32"`` -- what the reference does with ``dtype_f="float32"``: FP32
  frames, windows, FFT correlation (single-precision pocketfft standing in for
  cuFFT) and sub-pixel fit; smoothn in FP64 (gpu_process.py:842 builds
  ``SmoothnGPU`` without ``dtype_f`` so it takes the float64 default,
  gpu_smoothn.py:103), which promotes ``u, v`` to FP64 from pass 2 on
  (gpu_process.py:631-632).
* ``"fp64"`` -- the same algorithm with every array in FP64 (precision
  reference).

How each function is pinned (tests/test_oracle_pins.py):

=====================  =====================================================
function               pinned by
=====================  =====================================================
field_geometry         closed-form grid table for the documented workloads
get_stack              brute-force indexing
correlate_windows      direct O(F^4) circular correlation; integer-shift
                       known answer (peak at F/2 + d)
first_peak             np.argmax first-occurrence on planted ties
subpixel_gaussian      exact recovery of sampled Gaussians
rgi_linear             scipy.interpolate.RegularGridInterpolator
deform_frame           exact on affine intensity fields; brute force loop
median_validation      np.nanmedian over NaN-padded 3x3 sliding windows
spring_replacement     hand-solved 1-outlier and 2-outlier systems;
                       scipy.sparse.linalg.spsolve on the same matrix
smoothn_lambda         C^T diag(Lambda) C equals the square of the 2-D
                       Neumann Laplacian built from its stencil
smoothn                scipy.fft.dctn/idctn; fixed s equals the dense
                       solve of (I + s L^2) z = f; constant field invariance
gcv_score              the dense hat-matrix GCV with H = (I + s L^2)^-1, in
(+ weighted branch)    both branches; the chosen p vs the dense GCV minimum
piv (end to end)       RMS vs the known synthetic displacement
=====================  =====================================================

Parity unpinned (documented, no reference vectors exist and the reference
cannot run without CuPy/CUDA): cuFFT rounding, CuPy's FP32 reduction order,
cupyx ``spsolve`` behaviour on a singular spring system (all vectors
outliers; treated here as "unresolved -> reset", gpu_process.py:829-832), and
the exact point where SciPy's L-BFGS-B stops on GPU-evaluated GCV scores.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np
import scipy.fft as sfft
import scipy.sparse as sps
import scipy.sparse.linalg as spla
from scipy.interpolate import RegularGridInterpolator
from scipy.ndimage import distance_transform_edt
from scipy.optimize import fmin_l_bfgs_b

EPS_SUBPIXEL = 1e-7  # gpu_process.py:1011


# ---------------------------------------------------------------------------
# Geometry -- PIVFieldGPU (gpu_process.py:897-961) and the pass schedule
# (gpu_process.py:675-680).
# ---------------------------------------------------------------------------

@dataclass
class Geometry:
    frame_shape: Tuple[int, int]
    search_size: int
    overlap: int
    spacing: int
    field_shape: Tuple[int, int]
    offset: Tuple[int, int]  # (x_offset, y_offset)
    x: np.ndarray            # (m, n) float32 window-centre x
    y: np.ndarray            # (m, n) float32 window-centre y

    @property
    def x_grid(self):
        return self.x[0, :]

    @property
    def y_grid(self):
        return self.y[:, 0]

    @property
    def n_windows(self):
        return self.field_shape[0] * self.field_shape[1]


def field_geometry(frame_shape, search_size, overlap_ratio, center=True, dtype=np.float32) -> Geometry:
    """gpu_process.py:904-947 (window_size == search_size, no shrink)."""
    ht, wd = frame_shape
    overlap = int(overlap_ratio * search_size)                      # :907
    spacing = search_size - overlap                                 # :910
    m = (ht - search_size) // spacing + 1                           # :922-925
    n = (wd - search_size) // spacing + 1
    if center:                                                      # :927-938
        x_off = (wd - (spacing * (n - 1) + search_size)) // 2 + search_size % 2
        y_off = (ht - (spacing * (m - 1) + search_size)) // 2 + search_size % 2
    else:
        x_off = y_off = 0
    x1 = np.arange(n) * spacing + search_size / 2.0 + x_off          # :944
    y1 = np.arange(m) * spacing + search_size / 2.0 + y_off          # :945
    x, y = np.meshgrid(x1, y1)                                      # :946
    return Geometry((ht, wd), search_size, overlap, spacing, (m, n), (x_off, y_off),
                    x.astype(dtype), y.astype(dtype))


def pass_sizes(min_search_size, search_size_iters) -> List[int]:
    """gpu_process.py:675-680: pass i uses 2**(P-1-i) * min_search_size,
    repeated search_size_iters[i] times."""
    iters = (search_size_iters,) if isinstance(search_size_iters, int) else tuple(search_size_iters)
    out = []
    for i, it in enumerate(iters):
        out += [(2 ** (len(iters) - i - 1)) * min_search_size] * it
    return out


def fft_size(search_size, n_fft=2):
    """gpu_process.py:1099-1109."""
    return 2 ** int(math.ceil(math.log2(search_size * n_fft)))


# ---------------------------------------------------------------------------
# Extraction -- cuda_get_stack (gpu_process.py:1420-1449).
# ---------------------------------------------------------------------------

def get_stack(frame: np.ndarray, geom: Geometry) -> np.ndarray:
    """stack[k, i, j] = f[(k//n)*sp + y_off + i, (k%n)*sp + x_off + j]."""
    m, n = geom.field_shape
    ss, sp = geom.search_size, geom.spacing
    x_off, y_off = geom.offset
    view = np.lib.stride_tricks.sliding_window_view(frame, (ss, ss))
    rows = np.arange(m) * sp + y_off
    cols = np.arange(n) * sp + x_off
    return np.ascontiguousarray(view[rows][:, cols].reshape(m * n, ss, ss))


# ---------------------------------------------------------------------------
# Linear RGI with extrapolation -- cupyx RegularGridInterpolator(method
# "linear", bounds_error=False, fill_value=None) as used at
# gpu_process.py:737-740 and :1159-1162.  Indices: searchsorted(left)-1
# clipped to [0, n-2]; unclamped normalised distance, so points outside the
# grid extrapolate linearly from the edge cell.
# ---------------------------------------------------------------------------

def rgi_linear(y_grid, x_grid, values, yq, xq):
    yg = np.asarray(y_grid, np.float64)
    xg = np.asarray(x_grid, np.float64)
    vals = np.asarray(values, np.float64)
    yq = np.asarray(yq, np.float64)
    xq = np.asarray(xq, np.float64)
    iy = np.clip(np.searchsorted(yg, yq, side="left") - 1, 0, yg.size - 2)
    ix = np.clip(np.searchsorted(xg, xq, side="left") - 1, 0, xg.size - 2)
    ty = (yq - yg[iy]) / (yg[iy + 1] - yg[iy])
    tx = (xq - xg[ix]) / (xg[ix + 1] - xg[ix])
    return (vals[iy, ix] * (1 - ty) * (1 - tx) + vals[iy, ix + 1] * (1 - ty) * tx
            + vals[iy + 1, ix] * ty * (1 - tx) + vals[iy + 1, ix + 1] * ty * tx)


def interpolate_displacement(prev: Geometry, cur: Geometry, up, vp):
    """gpu_process.py:730-742: previous grid -> current grid, only when the
    search size changes."""
    if prev.search_size == cur.search_size:
        return up, vp
    yq, xq = cur.y, cur.x
    return (rgi_linear(prev.y_grid, prev.x_grid, up, yq, xq),
            rgi_linear(prev.y_grid, prev.x_grid, vp, yq, xq))


# ---------------------------------------------------------------------------
# Frame deformation -- deform_frame + cuda_interpolate_frame
# (gpu_process.py:1149-1192, 1508-1547).
# ---------------------------------------------------------------------------

def _f32(a):
    return np.asarray(a, dtype=np.float32)


def warp_frame(f, up_pix, vp_pix, dt, dtype=np.float32, mask=None):
    """Backward bilinear warp, clamped base index, unclamped fraction
    (gpu_process.py:1528-1546).  ``xi = x + up*dt`` is formed as one fused
    multiply-add (rounded once), the way NVRTC contracts it.  Masked pixels
    stay 0 (cuda_interpolate_masked_frame, :1461-1506)."""
    ht, wd = f.shape
    f = np.asarray(f, dtype)
    yy, xx = np.meshgrid(np.arange(ht, dtype=dtype), np.arange(wd, dtype=dtype), indexing="ij")
    dtf = dtype(dt)
    if dtype == np.float32:
        xi = (xx.astype(np.float64) + up_pix.astype(np.float64) * np.float64(dtf)).astype(np.float32)
        yi = (yy.astype(np.float64) + vp_pix.astype(np.float64) * np.float64(dtf)).astype(np.float32)
    else:
        xi = xx + up_pix * dtf
        yi = yy + vp_pix * dtf
    xl = np.floor(xi).astype(np.int64)
    yu = np.floor(yi).astype(np.int64)
    yu = np.clip(yu, 0, ht - 2)
    xl = np.clip(xl, 0, wd - 2)
    yd = yu + 1
    xr = xl + 1
    ful = f[yu, xl]
    fur = f[yu, xr]
    fdl = f[yd, xl]
    fdr = f[yd, xr]
    ty0 = (yi - yu.astype(dtype))
    ty1 = (yd.astype(dtype) - yi)
    tx0 = (xr.astype(dtype) - xi)
    tx1 = (xi - xl.astype(dtype))
    out = (ty0 * tx0 * fdl + ty0 * tx1 * fdr + ty1 * tx0 * ful + ty1 * tx1 * fur).astype(dtype)
    if mask is not None:
        out[mask] = 0
    return out


def deform_frames(frame_a, frame_b, geom: Geometry, up, vp, deforming_par=0.5, dtype=np.float32, mask=None):
    """gpu_process.py:1149-1173: predictor interpolated (FP64 RGI) from the
    current grid onto every pixel, cast to dtype_f, then frame A warped at
    ``dt-1`` and frame B at ``dt`` (0 < dt < 1); dt == 1 warps B only,
    dt == 0 warps A only."""
    ht, wd = geom.frame_shape
    yy, xx = np.meshgrid(np.arange(ht), np.arange(wd), indexing="ij")
    up_pix = rgi_linear(geom.y_grid, geom.x_grid, up, yy, xx).astype(dtype)
    vp_pix = rgi_linear(geom.y_grid, geom.x_grid, vp, yy, xx).astype(dtype)
    dt = deforming_par
    if 0 < dt < 1:
        return (warp_frame(frame_a, up_pix, vp_pix, dt - 1, dtype, mask),
                warp_frame(frame_b, up_pix, vp_pix, dt, dtype, mask))
    if dt == 1:
        return np.asarray(frame_a, dtype), warp_frame(frame_b, up_pix, vp_pix, 1, dtype, mask)
    return warp_frame(frame_a, up_pix, vp_pix, -1, dtype, mask), np.asarray(frame_b, dtype)


def field_mask(geom: Geometry, mask):
    """PIVFieldGPU.get_field_mask (gpu_process.py:949-951): the frame mask
    at the window centres, mask[int(y), int(x)] of the dtype_f coordinates."""
    if mask is None:
        return None
    return np.asarray(mask, bool)[geom.y.astype(int), geom.x.astype(int)]


def mask_frames(frame_a, frame_b, mask):
    """PIVGPU.mask_frames (gpu_process.py:682-687) on the dtype_f copies."""
    if mask is not None:
        frame_a = frame_a.copy()
        frame_b = frame_b.copy()
        frame_a[mask] = 0
        frame_b[mask] = 0
    return frame_a, frame_b


# ---------------------------------------------------------------------------
# Correlation -- normalize_intensity, zero padding, rfft2/irfft2, fftshift
# (gpu_process.py:1065, 1201-1246).
# ---------------------------------------------------------------------------

def correlate_windows(win_a, win_b, fsize, normalize=True, win_pad=0):
    """Returns the fftshift-ed (N, F, F) correlation planes.

    Normalisation is forced on whenever padding is nonzero
    (gpu_process.py:1065) and subtracts the per-window mean over ss x ss
    before zero padding (:1201-1205).  The planes are
    ``irfft2(conj(rfft2(a)) * rfft2(b))`` (:1237-1246), i.e. the circular
    cross-correlation ``c[d] = sum_x a[x] b[x+d]`` with zero lag at F/2 after
    the shift (:1218).
    """
    dtype = win_a.dtype
    n, ss, _ = win_a.shape
    pad = (fsize - ss) // 2
    if win_pad:
        # interrogation window inside the search window: A is zero outside
        # it and normalised over it only (gpu_process.py:1066-1081, 1198-1199)
        inner = (slice(None), slice(win_pad, ss - win_pad), slice(win_pad, ss - win_pad))
        a = np.zeros_like(win_a)
        a[inner] = win_a[inner]
        if pad != 0 or normalize:
            a[inner] = a[inner] - a[inner].mean(axis=(1, 2), keepdims=True, dtype=dtype)
            win_b = win_b - win_b.mean(axis=(1, 2), keepdims=True, dtype=dtype)
        win_a = a
    elif pad != 0 or normalize:
        win_a = win_a - win_a.mean(axis=(1, 2), keepdims=True, dtype=dtype)
        win_b = win_b - win_b.mean(axis=(1, 2), keepdims=True, dtype=dtype)
    if pad != 0:
        pa = np.zeros((n, fsize, fsize), dtype)
        pb = np.zeros((n, fsize, fsize), dtype)
        pa[:, pad:pad + ss, pad:pad + ss] = win_a
        pb[:, pad:pad + ss, pad:pad + ss] = win_b
    else:
        pa, pb = win_a, win_b
    fa = sfft.rfft2(pa, axes=(1, 2))
    fb = sfft.rfft2(pb, axes=(1, 2))
    corr = sfft.irfft2(np.conj(fa) * fb, s=(fsize, fsize), axes=(1, 2)).astype(dtype)
    return sfft.fftshift(corr, axes=(1, 2))


def first_peak(corr):
    """gpu_process.py:1248-1254: flattened argmax, first occurrence wins."""
    n, fh, fw = corr.shape
    idx = np.argmax(corr.reshape(n, -1), axis=1)
    return idx // fw, idx % fw


def subpixel_peak(corr, ip, jp, method="gaussian"):
    """gpu_process.py:1256-1294.  The fit runs only for non-border peaks
    (:1262) and, for the Gaussian method, only where all five samples are > 0
    (:1277); elsewhere the integer peak is kept.  The FP32 fraction is added
    to the integer index in FP64 and stored back into dtype_f."""
    dtype = corr.dtype.type
    n, fh, fw = corr.shape
    i_sp = ip.astype(dtype)
    j_sp = jp.astype(dtype)
    ok = (ip > 0) & (ip < fh - 1) & (jp > 0) & (jp < fw - 1)
    k = np.nonzero(ok)[0]
    ic, jc = ip[k], jp[k]
    c = corr[k, ic, jc]
    cd = corr[k, ic - 1, jc]
    cu = corr[k, ic + 1, jc]
    cl = corr[k, ic, jc - 1]
    cr = corr[k, ic, jc + 1]
    eps = dtype(EPS_SUBPIXEL)
    with np.errstate(divide="ignore", invalid="ignore"):
        if method == "gaussian":
            g = (c > 0) & (cd > 0) & (cu > 0) & (cl > 0) & (cr > 0)
            k, ic, jc = k[g], ic[g], jc[g]
            c, cd, cu, cl, cr = (np.log(t[g]) for t in (c, cd, cu, cl, cr))
            di = dtype(0.5) * (cd - cu) / (cd - dtype(2.0) * c + cu + eps)
            dj = dtype(0.5) * (cl - cr) / (cl - dtype(2.0) * c + cr + eps)
        elif method == "parabolic":
            di = dtype(0.5) * (cd - cu) / (cd - dtype(2.0) * c + cu + eps)
            dj = dtype(0.5) * (cl - cr) / (cl - dtype(2.0) * c + cr + eps)
        elif method == "centroid":
            di = (cu - cd) / (cd + c + cu + eps)
            dj = (cr - cl) / (cl + c + cr + eps)
        else:
            raise ValueError(method)
    i_sp[k] = (ic.astype(np.float64) + di.astype(np.float64)).astype(dtype)
    j_sp[k] = (jc.astype(np.float64) + dj.astype(np.float64)).astype(dtype)
    return i_sp, j_sp


def correlate_pass(frame_a, frame_b, geom: Geometry, n_fft=2, normalize=True,
                   subpixel_method="gaussian", dtype=np.float32, return_planes=False, window_size=None):
    """CorrelationGPU.__call__ (gpu_process.py:1021-1097) on already deformed
    frames: stack, normalise, correlate, peak, sub-pixel, displacement
    ``u = j_sp - F//2``, ``v = i_sp - F//2`` (:1296-1310)."""
    fsize = fft_size(geom.search_size, n_fft)
    wa = get_stack(np.asarray(frame_a, dtype), geom)
    wb = get_stack(np.asarray(frame_b, dtype), geom)
    win_pad = (geom.search_size - window_size) // 2 if window_size else 0
    corr = correlate_windows(wa, wb, fsize, normalize, win_pad)
    ip, jp = first_peak(corr)
    i_sp, j_sp = subpixel_peak(corr, ip, jp, subpixel_method)
    m, n = geom.field_shape
    u = (j_sp - dtype(fsize // 2)).reshape(m, n)
    v = (i_sp - dtype(fsize // 2)).reshape(m, n)
    out = {"u": u, "v": v, "i_peak": ip.reshape(m, n), "j_peak": jp.reshape(m, n)}
    if return_planes:
        out["corr"] = corr
    return out


# ---------------------------------------------------------------------------
# Median validation -- ValidationGPU.median_validation (gpu_validation.py:
# 136-141, 166-210; gpu_misc.py:22-50).
# ---------------------------------------------------------------------------

def neighbour_median(f, size=1):
    """NaN-padded (2*size+1)^2 neighbourhoods with the centre set to NaN,
    reduced by nanmedian (gpu_validation.py:166-205)."""
    ks = 2 * size + 1
    fp = np.pad(f, size, mode="constant", constant_values=np.nan)
    win = np.lib.stride_tricks.sliding_window_view(fp, (ks, ks)).copy()
    win[:, :, size, size] = np.nan
    m, n = f.shape
    with np.errstate(all="ignore"):
        import warnings
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            return np.nanmedian(win.reshape(m, n, ks * ks), axis=2).astype(f.dtype)


def median_validation(u, v, median_tol=2.0, size=1):
    """Flags ``|f - median| > tol`` OR-ed over u and v (gpu_validation.py:
    136-141); an absolute test in pixels."""
    mu = neighbour_median(u, size)
    mv = neighbour_median(v, size)
    with np.errstate(invalid="ignore"):
        return (np.abs(u - mu) > median_tol) | (np.abs(v - mv) > median_tol)


def median_residual_ratio(u, v, median_tol=2.0, size=1):
    """max(|u - med_u|, |v - med_v|) / tol, for excluding threshold ties."""
    mu = neighbour_median(u, size)
    mv = neighbour_median(v, size)
    return np.maximum(np.abs(u - mu), np.abs(v - mv)) / median_tol


# ---------------------------------------------------------------------------
# Signal-to-noise ratio -- CorrelationGPU.get_s2n_ratio / get_second_peak /
# get_correlation_energy (gpu_process.py:1312-1365; gpu_misc.py:52-85).
# ---------------------------------------------------------------------------

def s2n_ratio(corr, ip, jp, method="peak2peak", size=2):
    """Per-window S2N from the fftshift-ed planes (N, F, F).

    * peak2peak: the (2 size + 1)^2 block around the first peak is set to
      -inf (``fill_kernel``; CuPy wraps out-of-range fancy indices, so the
      block wraps around the plane edges) and the first-occurrence argmax of
      what is left is the second peak (:1318-1320, 1345-1354);
    * peak2mean: the mean of the values in (0, peak / 2] (:1322-1324, 1356-1360);
    * peak2energy: the mean of the positive values (:1326-1327).
    ``s2n = peak1 / peak2`` where peak1 > 0, +inf where additionally
    peak2 <= 0, 0 where peak1 <= 0 (:1329-1333); an empty mean is NaN and so
    is the ratio.
    """
    corr = np.asarray(corr)
    dtype = corr.dtype.type
    nw, fh, fw = corr.shape
    k = np.arange(nw)
    ip = np.asarray(ip).reshape(-1)
    jp = np.asarray(jp).reshape(-1)
    c1 = corr[k, ip, jp]
    with np.errstate(all="ignore"):
        if method == "peak2peak":
            work = corr.copy()
            sh = np.arange(-size, size + 1)
            ii = (ip[:, None, None] + sh[None, :, None]) % fh
            jj = (jp[:, None, None] + sh[None, None, :]) % fw
            work[k[:, None, None], ii, jj] = -np.inf
            i2, j2 = first_peak(work)
            c2 = work[k, i2, j2]
        else:
            c = np.where(corr > dtype(0.5) * c1[:, None, None], np.nan, corr) if method == "peak2mean" else corr
            c = np.where(c <= 0, np.nan, c)
            import warnings
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", RuntimeWarning)
                c2 = np.nanmean(c, axis=(1, 2)).astype(dtype)
        s2n = np.where(c1 > 0, c1 / c2, dtype(0))
        s2n = np.where((c2 <= 0) & (c1 > 0), dtype(np.inf), s2n)
    return s2n.astype(dtype)


# ---------------------------------------------------------------------------
# Full validation -- ValidationGPU.__call__ (gpu_validation.py:84-164,
# 166-227) with every test, and the replacement loop
# (gpu_validation.py:294-415; gpu_process.py:744-834).
# ---------------------------------------------------------------------------

def _stacks(f, size=1, centre_nan=True):
    ks = 2 * size + 1
    f = np.asarray(f)
    if not np.issubdtype(f.dtype, np.floating):
        f = f.astype(np.float64)
    fp = np.pad(f, size, mode="constant", constant_values=np.nan)
    win = np.lib.stride_tricks.sliding_window_view(fp, (ks, ks)).copy()
    if centre_nan:
        win[:, :, size, size] = np.nan
    m, n = np.shape(f)
    return win.reshape(m, n, ks * ks)


def _nan_reduce(fn, a):
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        return fn(a, axis=-1)


def neighbour_stats(f, size=1):
    """Median, MAD (median of |x - median|) and mean of each NaN-padded
    neighbourhood with the centre excluded (gpu_validation.py:166-205)."""
    st = _stacks(f, size)
    med = _nan_reduce(np.nanmedian, st)
    mad = _nan_reduce(np.nanmedian, np.abs(st - med[..., None]))
    mean = _nan_reduce(np.nanmean, st)
    return med, mad, mean


def validate(u, v, s2n=None, s2n_tol=None, median_tol=2.0, mad_tol=None, mean_tol=None, rms_tol=None, size=1,
             only=None):
    """OR of the enabled tests (gpu_validation.py:103-118), per field OR-ed
    over u and v: s2n < s2n_tol; |f - med| > median_tol; |f - med| >
    mad_tol * mad; |f - mean| > mean_tol; |f - mean(f)| > rms_tol * std(f)
    (global NaN-mean and population std, :143-148).  ``only``: positions to
    validate (the reference's mask complement, :119, 222-226); others are
    False.  With no test enabled, nothing is flagged."""
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    flags = np.zeros(u.shape, bool)
    with np.errstate(invalid="ignore"):
        if s2n_tol is not None and s2n is not None:
            flags |= np.asarray(s2n).reshape(u.shape) < s2n_tol
        if median_tol is not None or mad_tol is not None or mean_tol is not None:
            su, sv = neighbour_stats(u, size), neighbour_stats(v, size)
            if median_tol is not None:
                flags |= (np.abs(u - su[0]) > median_tol) | (np.abs(v - sv[0]) > median_tol)
            if mad_tol is not None:
                flags |= (np.abs(u - su[0]) > mad_tol * su[1]) | (np.abs(v - sv[0]) > mad_tol * sv[1])
            if mean_tol is not None:
                flags |= (np.abs(u - su[2]) > mean_tol) | (np.abs(v - sv[2]) > mean_tol)
        if rms_tol is not None:
            for f in (u, v):
                mu = _nan_reduce(np.nanmean, f.reshape(-1))
                sd = _nan_reduce(np.nanstd, f.reshape(-1))
                flags |= np.abs(f - mu) > rms_tol * sd
    if only is not None:
        flags &= only
    return flags


def neighbour_replacement(u, v, val, method="median", exclude_flagged=True, size=1):
    """ReplacementGPU median/mean replacement (gpu_validation.py:373-398):
    every flagged vector takes the nanmedian / nanmean of its neighbourhood
    (centre excluded), with the flagged vectors themselves excluded
    (``fill_value=NaN``) or included at their current values (``None``), all
    from the same input (simultaneous update)."""
    out = []
    for f in (u, v):
        f = np.asarray(f)
        g = np.where(val, f.dtype.type(np.nan), f) if exclude_flagged else f
        st = _stacks(g, size)[val]
        stat = _nan_reduce(np.nanmedian if method == "median" else np.nanmean, st)
        h = f.copy()
        h[val] = stat  # written into the field's own dtype, as the reference assigns (:328-329)
        out.append(h)
    return out[0], out[1]


def replace_outliers(u, v, val, method="spring", n_iters=1, revalidate=False, validate_fn=None, size=1):
    """PIVGPU.replace_outliers (gpu_process.py:788-834).  Returns (u, v,
    val): the replaced fields and the flags after the loop (updated only when
    revalidating).  Iteration 1 excludes the flagged vectors; without
    revalidation later iterations include them at their replaced values
    (:824-826); with it, unresolved (NaN) vectors are reset to the originals
    and only the previously flagged vectors are re-tested (:809-822).
    Remaining NaN vectors are reset at the end (:829-832)."""
    u = np.asarray(u)
    v = np.asarray(v)
    val = np.asarray(val, bool)
    if not val.any() or n_iters == 0:
        return u, v, val
    u0, v0 = u.copy(), v.copy()
    exclude = True
    last_val = val
    for _ in range(n_iters):
        if not val.any():
            break
        if method == "spring":
            u, v = spring_replacement(u, v, val, size)
        else:
            u, v = neighbour_replacement(u, v, val, method, exclude, size)
        last_val = val
        if revalidate:
            nan = (np.isnan(u) | np.isnan(v)) & val
            u = np.where(nan, u0, u)
            v = np.where(nan, v0, v)
            val = validate_fn(u, v, only=val) | nan
        else:
            exclude = False
    if not revalidate:
        nan = (np.isnan(u) | np.isnan(v)) & last_val
        u = np.where(nan, u0, u)
        v = np.where(nan, v0, v)
    return u, v, val


# ---------------------------------------------------------------------------
# Spring replacement -- ReplacementGPU.spring_replacement
# (gpu_validation.py:336-371, 381-415) and PIVGPU.replace_outliers
# (gpu_process.py:788-834).
# ---------------------------------------------------------------------------

def spring_system(val, f_list, size=1):
    """Builds the reference's link matrix and right-hand sides.

    Row k (outlier at (i, j)): ``x_k + c_k * sum_{l in N4(k) & O} x_l = rhs``
    where N4(k) are the four vectors ``size`` rows or columns away
    (gpu_validation.py:279-283), ``c_k`` = -1/4 interior, -1/3 in the
    ``size``-wide border, -1/2 in its four size x size corners, by *position*
    (:365-371), and ``rhs`` the NaN-mean over the in-bounds N4(k) with the
    outliers set to 0 (:342, 381-394).
    """
    m, n = val.shape
    s = size
    coef = np.full((m, n), -1.0 / 3.0)
    coef[s:-s, s:-s] = -0.25
    coef[:s, :s] = coef[:s, -s:] = coef[-s:, :s] = coef[-s:, -s:] = -0.5
    iv, jv = np.nonzero(val)
    nv = iv.size
    slot = -np.ones((m, n), np.int64)
    slot[iv, jv] = np.arange(nv)
    rows, cols, data = [], [], []
    for r in range(nv):
        rows.append(r); cols.append(r); data.append(1.0)
        for di, dj in ((0, -s), (0, s), (s, 0), (-s, 0)):
            a, b = iv[r] + di, jv[r] + dj
            if 0 <= a < m and 0 <= b < n and slot[a, b] >= 0:
                rows.append(r); cols.append(slot[a, b]); data.append(coef[iv[r], jv[r]])
    link = sps.csr_matrix((data, (rows, cols)), shape=(nv, nv))
    rhs = []
    for f in f_list:
        g = np.where(val, 0.0, np.asarray(f, np.float64))
        gp = np.pad(g, s, constant_values=np.nan)
        nb = np.stack([gp[s + iv, jv], gp[s + iv, 2 * s + jv], gp[2 * s + iv, s + jv], gp[iv, s + jv]], 0)
        import warnings
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            rhs.append(np.nanmean(nb, axis=0))
    return link, rhs, (iv, jv)


def spring_singular(val, size=1):
    """Outliers whose spring rows form a singular block.  Links ``size``
    apart join only vectors of one residue class (row mod size, column mod
    size), so the system splits into size^2 independent blocks; a block is
    singular when its whole class is flagged and no member sits where the
    position-based coefficient exceeds its in-bounds link count (rows in
    [m - size, size) or columns in [n - size, size), possible only in fields
    narrower than 2 size).  Such vectors have no solution; the reference's
    unresolved-vector reset (gpu_process.py:829-832) keeps their values."""
    val = np.asarray(val, bool)
    m, n = val.shape
    s = size
    out = np.zeros_like(val)
    for a in range(min(s, m)):
        for b in range(min(s, n)):
            if a >= m - s or b >= n - s:
                continue
            cls = val[a::s, b::s]
            if cls.all():
                out[a::s, b::s] = True
    return out


def spring_replacement(u, v, val, size=1):
    """Solves the spring system for both fields (direct sparse solve) and
    writes the results at the outlier positions.  A singular system leaves
    NaN, which the caller resets to the original values
    (gpu_process.py:829-832); with size 1 only a field of outliers is
    singular, with wider links also an outlier cluster that covers a whole
    residue class of the field (its rows and columns mod size), which
    :func:`spring_singular` detects."""
    u = np.array(u, copy=True)
    v = np.array(v, copy=True)
    n_vals = int(np.count_nonzero(val))
    if n_vals == 0:
        return u, v
    if n_vals == val.size:
        return u, v  # singular: unresolved everywhere -> reset to originals
    sing = spring_singular(val, size)
    if sing.any():
        val = val & ~sing
        if not val.any():
            return u, v
    link, rhs, (iv, jv) = spring_system(val, (u, v), size)
    for f, b in zip((u, v), rhs):
        x = spla.spsolve(link.tocsc(), b)
        x = np.where(np.isfinite(x), x, f[iv, jv])
        f[iv, jv] = x
    return u, v


# ---------------------------------------------------------------------------
# smoothn, unweighted automatic path -- SmoothnGPU.__call__
# (gpu_smoothn.py:123-231, 299-372).
# ---------------------------------------------------------------------------

def smoothn_p_bounds(shape, order=2):
    """gpu_smoothn.py:320-338."""
    h_min = 1e-3
    h_max = 1 - h_min
    rank = int(np.sum(np.array(shape) != 1))
    if order == 0:
        return math.log10(1 / h_max ** (1 / rank) - 1), math.log10(1 / h_min ** (1 / rank) - 1)
    if order == 1:
        return (math.log10((1 / (h_max ** (2 / rank)) - 1) / 4),
                math.log10((1 / (h_min ** (2 / rank)) - 1) / 4))
    p_min = math.log10((((1 + math.sqrt(1 + 8 * h_max ** (2 / rank))) / 4 / h_max ** (2 / rank)) ** 2 - 1) / 16)
    p_max = math.log10((((1 + math.sqrt(1 + 8 * h_min ** (2 / rank))) / 4 / h_min ** (2 / rank)) ** 2 - 1) / 16)
    return p_min, p_max


def smoothn_lambda(shape, order=2):
    """gpu_smoothn.py:340-351, raised to the smoothing order (:165)."""
    m, n = shape
    lam = (np.cos(np.pi * np.arange(m) / m)[:, None] + np.cos(np.pi * np.arange(n) / n)[None, :])
    return (2.0 * (2 - lam)) ** order


def dct2(z):
    """gpu_smoothn.py:299-310: orthonormal 2-D DCT-II via two 1-D passes."""
    return sfft.dctn(z, type=2, norm="ortho")


def idct2(z):
    return sfft.idctn(z, type=2, norm="ortho")


def gcv_score(p, f_dct, lam, n_fields):
    """gpu_smoothn.py:353-372, unweighted branch (w_mean > 0.9)."""
    p = float(np.asarray(p).reshape(-1)[0])
    gamma = 1.0 / (1.0 + 10.0 ** p * lam)
    rss = 0.0
    for fd in f_dct:
        rss += np.linalg.norm(fd * (gamma - 1)) ** 2
    tr_h = np.sum(gamma)
    size = lam.size
    return rss / size / (1 - tr_h / size) ** 2 / n_fields


def smoothn(fields: Sequence[np.ndarray], s=None, return_p=False, mask=None, tol_z=1e-3, max_iters=100,
            return_info=False):
    """Garcia's DCT-PLS smoother as ``SmoothnGPU.__call__`` runs it
    (gpu_smoothn.py:123-231), non-robust, order 2.

    Unweighted (no masked and no missing value): ``tol = 0`` so one
    iteration runs; with ``s=None`` the smoothing parameter is chosen by GCV:
    the best of 10 evenly spaced ``p = log10 s`` values, then
    ``fmin_l_bfgs_b(approx_grad, factr=10)`` within ``[p_min, p_max]``
    (:179-193).

    Weighted (``mask`` has a True element or a value is not finite,
    :253-269): missing values are filled from their nearest finite
    neighbour (:233-251), weights are 1 at finite unmasked values and 0
    elsewhere, z starts from a coarse DCT smoothing (:271-297), and the
    relaxed (factor 1.75) iteration runs until the relative change of z
    falls to ``tol_z`` or ``max_iters`` iterations, with GCV re-run at the
    iterations k where log2(k + 1) is an integer (L-BFGS-B from the current
    p after the first span, :176-193); masked values are 0 on return
    (:224-225)."""
    f = [np.array(a, np.float64) for a in fields]
    shape = f[0].shape
    if f[0].size < 2:
        return (list(f), None) if return_p else list(f)
    n_fields = len(f)
    size = f[0].size
    lam = smoothn_lambda(shape)
    mask = None if mask is None else np.asarray(mask, bool)
    finite = np.ones(shape, bool)
    for a in f:
        finite &= np.isfinite(a)
    missing = ~finite
    if missing.any():                                                        # :233-251
        idx = nearest_finite_index(missing)
        for a in f:
            a[missing] = a[idx[0], idx[1]][missing]
    w = finite.astype(np.float64)                                            # :262
    if mask is not None:
        w = np.where(mask, 0.0, w)                                           # :266-268
    weighted = bool(np.any(w != 1))
    n_finite = int(finite.sum())
    r = 1.0 + 0.75 * weighted                                                # :162
    if weighted:
        z = smoothn_init_z(f)                                                # :271-297
    else:
        z = [np.zeros(shape) for _ in f]
    p = None
    p_min, p_max = smoothn_p_bounds(shape)
    if s is not None:
        p = math.log10(s)
    w_mean = float(np.mean(w))
    k = 0
    n_iter = 0
    for k in range(max_iters):                                               # :173-214
        z0 = z
        f_dct = [dct2(w * (a - b) + b) for a, b in zip(f, z)]
        if s is None and float(np.log2(k + 1)) % 1 == 0:
            args = (f_dct, lam, n_fields, f, w, finite, w_mean, n_finite)
            if not p:
                p_span = np.arange(10) * (1 / 9) * (p_max - p_min) + p_min
                g = np.array([gcv_score_weighted(pi, *args) for pi in p_span])
                p = p_span[np.argmin(g)]
            pv, _, _ = fmin_l_bfgs_b(gcv_score_weighted, np.array([p]), fprime=None, factr=10, approx_grad=True,
                                     bounds=[(p_min, p_max)], args=args)
            p = float(pv[0])
        gamma = 1.0 / (1.0 + 10.0 ** p * lam)
        z = [r * idct2(gamma * fd) + (1 - r) * b for fd, b in zip(f_dct, z)]
        n_iter = k + 1
        if weighted:
            tol = (np.linalg.norm(np.concatenate([a - b for a, b in zip(z0, z)]))
                   / (np.linalg.norm(np.concatenate(z)) + 1e-6))
        else:
            tol = 0.0
        if tol <= tol_z:
            break
    if mask is not None:
        z = [np.where(mask, 0.0, a) for a in z]
    out = (z, p) if return_p else z
    if return_info:
        return out, {"iterations": n_iter, "weighted": weighted, "w_mean": w_mean}
    return out


def nearest_finite_index(missing):
    """Index of the nearest non-missing element (unit spacing; the
    reference calls scipy's exact Euclidean feature transform before
    ``init_spacing`` runs, gpu_smoothn.py:244-251, 155)."""
    return distance_transform_edt(missing, return_distances=False, return_indices=True)


def smoothn_init_z(f):
    """gpu_smoothn.py:283-297: DCT, zero the coefficients whose every index
    is at least ceil(shape / 10) + 1, inverse DCT."""
    z = []
    for a in f:
        d = dct2(a)
        n0, n1 = (int(math.ceil(s / 10)) + 1 for s in d.shape)
        d[n0:, n1:] = 0.0
        z.append(idct2(d))
    return z


def gcv_score_weighted(p, f_dct, lam, n_fields, f, w, finite, w_mean, n_finite):
    """gpu_smoothn.py:353-372: with w_mean > 0.9 the residual comes from the
    spectrum alone; otherwise from the weighted data misfit at finite
    points.  Normalised by the finite count."""
    p = float(np.asarray(p).reshape(-1)[0])
    gamma = 1.0 / (1.0 + 10.0 ** p * lam)
    rss = 0.0
    if w_mean > 0.9:
        for fd in f_dct:
            rss += np.linalg.norm(fd * (gamma - 1)) ** 2
    else:
        for a, fd in zip(f, f_dct):
            fh = idct2(gamma * fd)
            rss += np.linalg.norm(np.sqrt(w[finite]) * (a[finite] - fh[finite])) ** 2
    tr_h = np.sum(gamma)
    return rss / n_finite / (1 - tr_h / lam.size) ** 2 / n_fields


def gcv_exact_minimiser(fields):
    """Diagnostic: the GCV minimiser located by dense sampling + bounded
    Brent refinement, independent of L-BFGS-B's stopping rule."""
    from scipy.optimize import minimize_scalar
    f = [np.asarray(a, np.float64) for a in fields]
    lam = smoothn_lambda(f[0].shape)
    f_dct = [dct2(a) for a in f]
    p_min, p_max = smoothn_p_bounds(f[0].shape)
    grid = np.linspace(p_min, p_max, 2001)
    g = np.array([gcv_score(p, f_dct, lam, len(f)) for p in grid])
    k = int(np.argmin(g))
    lo, hi = grid[max(k - 1, 0)], grid[min(k + 1, grid.size - 1)]
    r = minimize_scalar(lambda p: gcv_score(p, f_dct, lam, len(f)), bounds=(lo, hi), method="bounded",
                        options={"xatol": 1e-12})
    return float(r.x)


# ---------------------------------------------------------------------------
# The pass loop -- PIVGPU.__call__ (gpu_process.py:564-647).
# ---------------------------------------------------------------------------

@dataclass
class OracleSettings:
    min_search_size: int
    search_size_iters: Tuple[int, ...] = (1,)
    overlap_ratio: float = 0.5
    center: bool = True
    n_fft: int = 2
    deforming_par: float = 0.5
    subpixel_method: str = "gaussian"
    normalize: bool = True
    median_tol: Optional[float] = 2.0
    num_replacing_iters: int = 1
    smoothing_par: Optional[float] = None
    s2n_method: str = "peak2peak"
    s2n_size: int = 2
    s2n_tol: Optional[float] = None
    mad_tol: Optional[float] = None
    mean_tol: Optional[float] = None
    rms_tol: Optional[float] = None
    replacing_method: str = "spring"
    revalidate: bool = False
    validation_size: int = 1
    replacing_size: int = 1
    shrink_ratio: float = 1.0
    dt: float = 1.0
    scaling_par: float = 1.0
    mask: Optional[np.ndarray] = None


@dataclass
class PassState:
    geom: Geometry
    up: Optional[np.ndarray]
    vp: Optional[np.ndarray]
    corr: dict
    u_pre: np.ndarray        # after residual + predictor, before validation
    v_pre: np.ndarray
    val: np.ndarray          # flags before replacement
    u_rep: np.ndarray        # after replacement
    v_rep: np.ndarray
    outliers: Optional[np.ndarray]
    u_out: np.ndarray        # after smoothing (non-final) or replacement
    v_out: np.ndarray
    smooth_p: Optional[float] = None


def shrunk_window(search_size, shrink_ratio):
    """gpu_process.py:654-662: the first pass's interrogation window."""
    ws = int(search_size * shrink_ratio)
    ws = ws if ws % 2 == 0 else ws + 1
    return max(ws, 8)


def run_pass(frame_a, frame_b, settings: OracleSettings, geom: Geometry, prev_geom: Optional[Geometry],
             u_prev, v_prev, is_last: bool, mode="faithful32", window_size=None) -> PassState:
    """One iteration of the main loop (gpu_process.py:596-639)."""
    dtype = np.float32 if mode == "faithful32" else np.float64
    fa = np.asarray(frame_a, dtype)
    fb = np.asarray(frame_b, dtype)
    fmask = field_mask(geom, settings.mask)
    up = vp = None
    if prev_geom is not None:
        up, vp = interpolate_displacement(prev_geom, geom, u_prev, v_prev)   # :620-625
        fa, fb = deform_frames(fa, fb, geom, up, vp, settings.deforming_par, dtype,
                               None if settings.mask is None else np.asarray(settings.mask, bool))
    corr = correlate_pass(fa, fb, geom, settings.n_fft, settings.normalize, settings.subpixel_method, dtype,
                          return_planes=settings.s2n_tol is not None, window_size=window_size)
    u, v = corr["u"], corr["v"]
    if fmask is not None:                                                    # :1296-1310
        u = np.where(fmask, np.nan, u).astype(u.dtype)
        v = np.where(fmask, np.nan, v).astype(v.dtype)
    if up is not None:                                                       # :631-632
        u, v = u + up, v + vp
    u_pre, v_pre = u, v
    s2n = None
    if settings.s2n_tol is not None:                                         # :755
        s2n = s2n_ratio(corr["corr"], corr["i_peak"], corr["j_peak"], settings.s2n_method,
                        settings.s2n_size).reshape(u.shape)
        if fmask is not None:                                                # :1338-1343
            s2n = np.where(fmask, 0.0, s2n).astype(s2n.dtype)
        corr["s2n"] = s2n
    keep = None if fmask is None else ~fmask                                 # gpu_validation.py:207-210

    def validate_fn(a, b, only=None):                                        # :758-769
        o = keep if only is None else (only if keep is None else only & keep)
        return validate(a, b, s2n, settings.s2n_tol, settings.median_tol, settings.mad_tol, settings.mean_tol,
                        settings.rms_tol, size=settings.validation_size, only=o)

    val = validate_fn(u, v)
    u_rep, v_rep, _ = replace_outliers(u, v, val, settings.replacing_method, settings.num_replacing_iters,
                                       settings.revalidate, validate_fn, settings.replacing_size)  # :788-834
    outliers = None
    if is_last:
        outliers = validate_fn(u_rep, v_rep)                                 # :783-784
    p = None
    if not is_last:                                                          # :637-639, Q2
        (u_out, v_out), p = smoothn([u_rep, v_rep], s=settings.smoothing_par, return_p=True, mask=fmask)
    else:
        u_out, v_out = u_rep, v_rep
    return PassState(geom, up, vp, corr, u_pre, v_pre, val, u_rep, v_rep, outliers, u_out, v_out, p)


def piv(frame_a, frame_b, settings: OracleSettings, mode="faithful32", keep_passes=False):
    """Full multi-pass run.  Returns dict with x, y, u, v, val_locations
    (last pass, before replacement) and outliers (after replacement)."""
    frame_shape = tuple(int(s) for s in np.shape(frame_a))
    sizes = pass_sizes(settings.min_search_size, settings.search_size_iters)
    geoms = [field_geometry(frame_shape, ss, settings.overlap_ratio, settings.center) for ss in sizes]
    u = v = None
    passes = []
    if settings.mask is not None:                                            # :586, 682-687
        dtype = np.float32 if mode == "faithful32" else np.float64
        frame_a, frame_b = mask_frames(np.asarray(frame_a, dtype), np.asarray(frame_b, dtype),
                                       np.asarray(settings.mask, bool))
    for k, g in enumerate(geoms):
        ws = shrunk_window(g.search_size, settings.shrink_ratio) if k == 0 and settings.shrink_ratio != 1 else None
        st = run_pass(frame_a, frame_b, settings, g, geoms[k - 1] if k > 0 else None, u, v,
                      k == len(geoms) - 1, mode, window_size=ws)
        u, v = st.u_out, st.v_out
        if keep_passes:
            passes.append(st)
        last = st
    if settings.dt != 1 or settings.scaling_par != 1:                        # :642-643, 847-849
        u = u * settings.scaling_par / settings.dt
        v = v * settings.scaling_par / settings.dt
    g = geoms[-1]
    out = {"x": g.x * settings.scaling_par, "y": g.y * settings.scaling_par, "u": u, "v": v,
           "val_locations": last.val, "outliers": last.outliers}
    if keep_passes:
        out["passes"] = passes
    return out
