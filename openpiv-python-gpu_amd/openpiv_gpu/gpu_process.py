"""PIV image analysis on an AMD MI355X -- drop-in for ``openpiv_gpu.gpu_process``.

Public surface (same names, keyword arguments, defaults and outputs as the
reference's gpu_process.py):

* :class:`piv_gpu` -- user wrapper with argument checks, NumPy in / NumPy out
  (reference gpu_process.py:41-352);
* :class:`PIVGPU` -- the iterative multigrid loop, device-array outputs
  (gpu_process.py:354-867);
* :class:`PIVFieldGPU` -- closed-form window geometry (gpu_process.py:869-961).

Every pass of the loop runs in hand-written HIP kernels for gfx950 behind the
C ABI in include/openpiv_hip.h (window deformation, fused FFT correlation,
peak + sub-pixel fit, median validation, spring replacement and smoothn, all
on the device, optionally captured in one hipGraph).  Settings the native path
does not implement raise ``NotImplementedError`` naming the option; nothing
falls back to the CPU.
"""

from __future__ import annotations

import ctypes
from math import ceil, log2, prod

import numpy as np

from . import DTYPE_f, DTYPE_i
from . import _native
from .gpu_validation import (ALLOWED_REPLACING_METHODS, MAD_TOL, MEAN_TOL, MEDIAN_TOL, REPLACING_METHOD,
                             REPLACING_SIZE, RMS_TOL, S2N_TOL, VALIDATION_SIZE)
from .gpu_smoothn import SMOOTHING_PAR

# Default settings (gpu_process.py:19-35).
SEARCH_SIZE_ITERS = 1
OVERLAP_RATIO = 0.5
SHRINK_RATIO = 1
CENTER = True
DEFORMING_ORDER = 1
NORMALIZE = True
SUBPIXEL_METHOD = "gaussian"
N_FFT = 2
DEFORMING_PAR = 0.5
BATCH_SIZE = "full"
S2N_METHOD = "peak2peak"
S2N_SIZE = 2
NUM_REPLACING_ITERS = 1
REVALIDATE = False
SMOOTH = True

# Allowed settings (gpu_process.py:37-39).
ALLOWED_SUBPIXEL_METHODS = {"gaussian", "parabolic", "centroid"}
ALLOWED_S2N_METHODS = {"peak2peak", "peak2mean", "peak2energy"}

_DEFAULTS = dict(search_size_iters=SEARCH_SIZE_ITERS, overlap_ratio=OVERLAP_RATIO, shrink_ratio=SHRINK_RATIO,
                 center=CENTER, deforming_order=DEFORMING_ORDER, normalize=NORMALIZE,
                 subpixel_method=SUBPIXEL_METHOD, n_fft=N_FFT, deforming_par=DEFORMING_PAR, batch_size=BATCH_SIZE,
                 s2n_method=S2N_METHOD, s2n_size=S2N_SIZE, validation_size=VALIDATION_SIZE, s2n_tol=S2N_TOL,
                 median_tol=MEDIAN_TOL, mad_tol=MAD_TOL, mean_tol=MEAN_TOL, rms_tol=RMS_TOL,
                 num_replacing_iters=NUM_REPLACING_ITERS, replacing_method=REPLACING_METHOD,
                 replacing_size=REPLACING_SIZE, revalidate=REVALIDATE, smooth=SMOOTH, smoothing_par=SMOOTHING_PAR,
                 dt=1, scaling_par=1, mask=None)

_KWARGS = tuple(_DEFAULTS) + ("dtype_f",)


def _is_real(x):
    return isinstance(x, (int, float)) and not isinstance(x, bool)


def _tuple_of(value, n, scalar_test):
    """Broadcasts a scalar setting to a per-size tuple, as PIVGPU.__init__
    does (gpu_process.py:497-546)."""
    return (value,) * n if scalar_test(value) else value


class HostArray(np.ndarray):
    """NumPy array with a CuPy-style ``.get()`` (returns a plain ndarray), for
    values the reference hands out as device arrays but that live on the host
    here (grid coordinates)."""

    def get(self):
        return np.asarray(self).copy()


def _host(a):
    return np.asarray(a).view(HostArray)


class DeviceArray:
    """Device-resident result (``PIVGPU`` outputs); ``.get()`` copies to the
    host like a CuPy array.  ``.buffer`` is the underlying device allocation
    (``.buffer.ptr`` its address)."""

    def __init__(self, buffer, dtype):
        self.buffer = buffer
        self.dtype = np.dtype(dtype)

    @property
    def shape(self):
        return tuple(self.buffer.shape)

    def get(self):
        return self.buffer.get().astype(self.dtype, copy=False)

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


# ---------------------------------------------------------------------------
# Geometry
# ---------------------------------------------------------------------------

class PIVFieldGPU:
    """Geometric information of a PIV field (gpu_process.py:869-961).

    Parameters
    ----------
    frame_shape : tuple
        Shape of the frames, (ht, wd).
    window_size : int
        Size of the interrogation window.
    search_size : int or None, optional
        Size of the search window.
    overlap_ratio : float, optional
        Ratio of the overlap (between 0 and 1) between the windows.
    center : bool, optional
        Whether to center the velocity field with respect to the frame edges.
    mask : ndarray or None, optional
        2D array with True values indicating the masked locations.
    dtype_f : str, optional
        Float data type.
    """

    def __init__(self, frame_shape, window_size, search_size=None, overlap_ratio=OVERLAP_RATIO, center=CENTER,
                 mask=None, dtype_f=np.float32):
        self.frame_shape = frame_shape
        self.window_size = window_size
        self.search_size = search_size if search_size is not None else window_size
        self.overlap = int(overlap_ratio * self.search_size)
        self.pad_ht = self.pad_wd = (self.search_size - self.window_size) // 2
        self.pad_shape = (self.pad_ht, self.pad_wd)
        self.spacing = self.search_size - self.overlap
        ht, wd = frame_shape
        self.field_shape = ((ht - self.search_size) // self.spacing + 1, (wd - self.search_size) // self.spacing + 1)
        self.n_wins = prod(self.field_shape)
        self.dtype_f = dtype_f
        self.center = center
        m, n = self.field_shape
        if center:
            x_off = (wd - (self.spacing * (n - 1) + self.search_size)) // 2 + self.search_size % 2
            y_off = (ht - (self.spacing * (m - 1) + self.search_size)) // 2 + self.search_size % 2
        else:
            x_off = y_off = 0
        self.offset = (x_off, y_off)
        x = np.arange(n) * self.spacing + self.search_size / 2.0 + x_off
        y = np.arange(m) * self.spacing + self.search_size / 2.0 + y_off
        xx, yy = np.meshgrid(x, y)
        self.x = _host(xx.astype(dtype_f))
        self.y = _host(yy.astype(dtype_f))
        self.x_grid = self.x[0, :]
        self.y_grid = self.y[:, 0]
        self.mask = mask
        self.field_mask = None if mask is None else _host(mask[self.y.astype(int), self.x.astype(int)])

    @property
    def coords(self):
        "Returns the coordinates of the velocity field."
        return self.x, self.y

    @property
    def grid_coords(self):
        "Returns the grid coordinates of the velocity field."
        return self.x_grid, self.y_grid


# ---------------------------------------------------------------------------
# The multi-pass loop
# ---------------------------------------------------------------------------

class PIVGPU:
    """Iterative multigrid PIV (gpu_process.py:354-867) on the MI355X.

    Same parameters and defaults as the reference; see :class:`piv_gpu`.
    Processing is FP32 (``dtype_f`` omitted or ``"float32"``).  Multi-pass
    results come back as float64 arrays, like the reference whose FP64 smoothn
    promotes ``u, v`` from the second pass on.
    """

    def __init__(self, frame_shape, min_search_size, search_size_iters=SEARCH_SIZE_ITERS,
                 overlap_ratio=OVERLAP_RATIO, shrink_ratio=SHRINK_RATIO, center=CENTER,
                 deforming_order=DEFORMING_ORDER, normalize=NORMALIZE, subpixel_method=SUBPIXEL_METHOD, n_fft=N_FFT,
                 deforming_par=DEFORMING_PAR, batch_size=BATCH_SIZE, s2n_method=S2N_METHOD, s2n_size=S2N_SIZE,
                 validation_size=VALIDATION_SIZE, s2n_tol=S2N_TOL, median_tol=MEDIAN_TOL, mad_tol=MAD_TOL,
                 mean_tol=MEAN_TOL, rms_tol=RMS_TOL, num_replacing_iters=NUM_REPLACING_ITERS,
                 replacing_method=REPLACING_METHOD, replacing_size=REPLACING_SIZE, revalidate=REVALIDATE,
                 smooth=SMOOTH, smoothing_par=SMOOTHING_PAR, dt=1, scaling_par=1, mask=None, dtype_f="float32",
                 device=0, use_graph=True, predictor=None, max_reject_fraction=0.05, predictor_pass=-1):
        # gpu_process.py:552: anything but the string "float32" selects FP64
        # processing (the default here is "float32")
        self.fp64 = not (isinstance(dtype_f, str) and dtype_f == "float32")
        if self.fp64 and predictor is not None:
            raise NotImplementedError("the learned predictor runs FP32 processing only")
        if mask is not None:
            mask = np.asarray(mask).astype(bool)                # gpu_process.py:548-549
            if mask.shape != tuple(frame_shape):
                raise ValueError(f"mask has shape {mask.shape}, the frames {tuple(frame_shape)}")
            if predictor is not None:
                raise NotImplementedError("a learned predictor cannot be combined with a mask")
        self.min_search_size = min_search_size
        self.ss_iters = (search_size_iters,) if isinstance(search_size_iters, int) else tuple(search_size_iters)
        self.n_passes = len(self.ss_iters)
        self.frame_shape = tuple(frame_shape)
        n = self.n_passes
        self.overlap_ratio = _tuple_of(overlap_ratio, n, lambda v: v == 0 or isinstance(v, float))
        self.shrink_ratio = shrink_ratio
        self.center = center
        self.deforming_order = _tuple_of(deforming_order, n, lambda v: isinstance(v, int))
        self.is_normalized = _tuple_of(normalize, n, lambda v: isinstance(v, bool))
        self.subpixel_method = _tuple_of(subpixel_method, n, lambda v: isinstance(v, str))
        self.n_fft = _tuple_of(n_fft, n, lambda v: isinstance(v, int))
        self.deforming_par = _tuple_of(deforming_par, n, lambda v: v == 0 or v == 1 or isinstance(v, float))
        self.batch_size = _tuple_of(batch_size, n, lambda v: isinstance(v, int) or v == "full")
        self.s2n_method = _tuple_of(s2n_method, n, lambda v: isinstance(v, str))
        self.s2n_size = _tuple_of(s2n_size, n, lambda v: isinstance(v, int))
        self.validation_size = _tuple_of(validation_size, n, lambda v: isinstance(v, int))
        tol_scalar = lambda v: v is None or _is_real(v)
        self.s2n_tol = _tuple_of(s2n_tol, n, tol_scalar)
        self.median_tol = _tuple_of(median_tol, n, tol_scalar)
        self.mad_tol = _tuple_of(mad_tol, n, tol_scalar)
        self.mean_tol = _tuple_of(mean_tol, n, tol_scalar)
        self.rms_tol = _tuple_of(rms_tol, n, tol_scalar)
        self.num_replacing_iters = _tuple_of(num_replacing_iters, n, lambda v: isinstance(v, int))
        self.replacing_method = _tuple_of(replacing_method, n, lambda v: isinstance(v, str))
        self.replacing_size = _tuple_of(replacing_size, n, lambda v: v is None or isinstance(v, int))
        self.is_revalidated = _tuple_of(revalidate, n, lambda v: isinstance(v, bool))
        self.is_smoothed = _tuple_of(smooth, n, lambda v: isinstance(v, bool))
        self.smoothing_par = _tuple_of(smoothing_par, n, tol_scalar)
        self.dt = dt
        self.scaling_par = scaling_par
        self.mask = mask
        self.dtype_f = np.float64 if self.fp64 else np.float32
        self.device = device
        self.use_graph = bool(use_graph)
        # learned coarse predictor (openpiv_gpu.learned.CoarseFlowNet): replaces
        # every pass but the last; a pair whose last-pass median test rejects
        # more than max_reject_fraction of the vectors reruns the full schedule
        self.predictor = predictor
        self.max_reject_fraction = float(max_reject_fraction)
        self.predictor_pass = int(predictor_pass)  # first pass fed by the network (-1: the last)

        for i in range(n):
            if self.validation_size[i] > _native.MAX_NEIGHBOURHOOD_SIZE:
                raise NotImplementedError(f"validation_size above {_native.MAX_NEIGHBOURHOOD_SIZE}")
            if (self.replacing_size[i] or 1) > _native.MAX_NEIGHBOURHOOD_SIZE:
                raise NotImplementedError(f"replacing_size above {_native.MAX_NEIGHBOURHOOD_SIZE}")

        self.piv_fields = None
        self.init_val_locations = self.val_locations = None
        self._plan = None
        self._last = None
        self.init_piv_fields()

    # -- geometry / schedule -------------------------------------------------
    def get_geom_settings(self):
        """Search size and overlap ratio of each pass (gpu_process.py:675-680)."""
        for i, ss_iters in enumerate(self.ss_iters):
            for _ in range(ss_iters):
                yield (2 ** (len(self.ss_iters) - i - 1)) * self.min_search_size, self.overlap_ratio[i]

    def _size_index(self):
        for i, ss_iters in enumerate(self.ss_iters):
            for _ in range(ss_iters):
                yield i

    def init_piv_fields(self):
        """Creates field objects for all iterations (gpu_process.py:649-673):
        with shrink_ratio != 1 the first pass correlates an interrogation
        window of even(int(ss * shrink_ratio)) >= 8 px inside its search window."""
        self.piv_fields = []
        for k, (ss, r) in enumerate(self.get_geom_settings()):
            ws = ss
            if k == 0 and self.shrink_ratio != 1:
                ws = int(ss * self.shrink_ratio)
                ws = ws if ws % 2 == 0 else ws + 1
                ws = max(ws, 8)
            self.piv_fields.append(PIVFieldGPU(self.frame_shape, window_size=ws, search_size=ss, overlap_ratio=r,
                                               center=self.center, mask=self.mask, dtype_f=self.dtype_f))

    def _plan_desc(self, in_dtype: int) -> _native.PlanDesc:
        fields = self.piv_fields
        if len(fields) > _native.MAX_PASSES:
            raise NotImplementedError(f"at most {_native.MAX_PASSES} passes")
        d = _native.PlanDesc()
        d.height, d.width = self.frame_shape
        d.in_dtype = in_dtype
        d.n_passes = len(fields)
        for k, (f, i) in enumerate(zip(fields, self._size_index())):
            fft = 2 ** ceil(log2(f.search_size * self.n_fft[i]))
            if fft > _native.MAX_FFT_SIZE:
                raise NotImplementedError(f"FFT size {fft} (search_size {f.search_size}, n_fft {self.n_fft[i]}): "
                                          f"the correlation kernels cover FFT sizes 16-{_native.MAX_FFT_SIZE}")
            if (fft - f.search_size) % 2:
                raise NotImplementedError("odd padding: search_size must be even")
            p = d.passes[k]
            p.search_size = f.search_size
            if f.window_size > f.search_size or (f.search_size - f.window_size) % 2:
                raise NotImplementedError("window_size must be an even size not above search_size")
            p.window_size = f.window_size if f.window_size != f.search_size else 0
            p.spacing = f.spacing
            p.x_offset, p.y_offset = f.offset
            p.field_m, p.field_n = f.field_shape
            p.fft_size = fft
            p.normalize = int(bool(self.is_normalized[i]) or fft != f.search_size)   # gpu_process.py:1065
            p.subpixel = _native.SUBPIXEL_CODES[self.subpixel_method[i]]
            p.deforming_par = float(self.deforming_par[i])
            tol = lambda t: float(t) if t is not None else float("nan")
            p.median_tol = tol(self.median_tol[i])
            p.s2n_tol = tol(self.s2n_tol[i])
            p.mad_tol = tol(self.mad_tol[i])
            p.mean_tol = tol(self.mean_tol[i])
            p.rms_tol = tol(self.rms_tol[i])
            p.validate = int(any(t is not None for t in (self.median_tol[i], self.s2n_tol[i], self.mad_tol[i],
                                                         self.mean_tol[i], self.rms_tol[i])))
            p.s2n_method = _native.S2N_CODES[self.s2n_method[i]]
            p.s2n_size = int(self.s2n_size[i])
            p.replacing_method = _native.REPLACING_CODES[self.replacing_method[i]]
            p.revalidate = int(bool(self.is_revalidated[i]))
            p.num_replacing_iters = int(self.num_replacing_iters[i])
            # gpu_validation.py:74, 273: the kernel half-width may not exceed
            # the field
            rsize = self.replacing_size[i] if self.replacing_size[i] is not None else 1
            for name, size in (("validation_size", self.validation_size[i]), ("replacing_size", rsize)):
                assert all(size <= item for item in f.field_shape), "size cannot exceed field_shape."
                setattr(p, name, int(size))
            if p.num_replacing_iters > _native.MAX_REPLACING_ITERS:
                raise NotImplementedError(f"num_replacing_iters above {_native.MAX_REPLACING_ITERS}")
            # gpu_process.py:841 tests the whole tuple, so every non-final pass
            # is smoothed whatever the per-pass flag says.
            p.smooth = int(bool(self.is_smoothed))
            sp = self.smoothing_par[i]
            p.smoothing_par = float(sp) if sp is not None else 0.0
        scale = 1.0
        if self.dt != 1 or self.scaling_par != 1:
            scale = float(self.scaling_par) / float(self.dt)
        d.out_scale = scale
        d.use_graph = int(self.use_graph)
        d.precision = _native.PIV_FP64 if self.fp64 else _native.PIV_FP32
        d.scaling_par = float(self.scaling_par)
        d.dt = float(self.dt)
        return d

    def _get_plan(self, in_dtype: int):
        if self._plan is None or self._plan.desc.in_dtype != in_dtype:
            # the plan is kept only once the mask and the predictor are set:
            # a caller who catches an error and calls again gets a fresh plan
            plan = _native.Plan(self._plan_desc(in_dtype), self.device)
            if self.mask is not None:
                plan.set_mask(self.mask)
            if self.predictor is not None:
                plan.set_predictor(self.predictor, self.max_reject_fraction, self.predictor_pass)
            self._plan = plan
        return self._plan

    @property
    def last_fallback(self):
        """Whether the last pair fell back to the full schedule (learned
        predictor only), and the last pass's rejected-vector count."""
        return self._plan.last_fallback() if self._plan is not None else (False, 0)

    # -- the pass loop ---------------------------------------------------------
    def _prepare_frame(self, frame):
        # the library reads uint8 / uint16 / float32 (and float64 when processing
        # in FP64) frames directly and converts exactly; anything else is cast
        # like gpu_process.py:582-583 (astype(dtype_f))
        frame = np.asarray(frame)
        raw = (np.uint8, np.uint16, np.float32) + ((np.float64,) if self.fp64 else ())
        if frame.dtype not in raw:
            frame = frame.astype(self.dtype_f)
        return np.ascontiguousarray(frame)

    def _run(self, frame_a, frame_b, outputs_on_device):
        fa = self._prepare_frame(frame_a)
        fb = self._prepare_frame(frame_b)
        if fa.dtype != fb.dtype:
            fa = fa.astype(self.dtype_f)
            fb = fb.astype(self.dtype_f)
        plan = self._get_plan(_native.in_dtype_code(fa.dtype))
        m, n = self.piv_fields[-1].field_shape
        fd = self.dtype_f
        if outputs_on_device:
            outs = [_native.DeviceBuffer((m, n), fd), _native.DeviceBuffer((m, n), fd),
                    _native.DeviceBuffer((m, n), np.uint8), _native.DeviceBuffer((m, n), np.uint8)]
            u, v, vi, vf = outs
            ptrs = [b.value for b in outs]
        else:
            u = np.empty((m, n), fd)
            v = np.empty((m, n), fd)
            vi = np.empty((m, n), np.uint8)
            vf = np.empty((m, n), np.uint8)
            ptrs = [a.ctypes.data for a in (u, v, vi, vf)]
        plan.run(fa.ctypes.data, fb.ctypes.data, False, *ptrs, outputs_on_device)
        return u, v, vi, vf

    @property
    def _out_dtype(self):
        # FP32 processing: one pass FP32, more FP64 through smoothn's promotion
        # (gpu_process.py:631-632, 842); FP64 processing: FP64
        return np.float32 if len(self.piv_fields) == 1 and not self.fp64 else np.float64

    def __call__(self, frame_a, frame_b):
        """Computes the velocity field from an image pair; returns device
        arrays ``(u, v)`` (gpu_process.py:564-647)."""
        u, v, vi, vf = self._run(frame_a, frame_b, outputs_on_device=True)
        self._store_flags(vi, vf)
        dt = self._out_dtype
        return DeviceArray(u, dt), DeviceArray(v, dt)

    def _store_flags(self, vi, vf):
        # the last pass belongs to the last size group
        if self.median_tol[len(self.ss_iters) - 1] is not None:
            self.init_val_locations = vi
            self.val_locations = vf
        else:
            self.init_val_locations = self.val_locations = None

    def run_host(self, frame_a, frame_b):
        """Host-to-host run used by :class:`piv_gpu`: NumPy in, NumPy out."""
        u, v, vi, vf = self._run(frame_a, frame_b, outputs_on_device=False)
        self._store_flags(vi, vf)
        dt = self._out_dtype
        return u.astype(dt), v.astype(dt)

    # -- accessors -----------------------------------------------------------------
    @property
    def coords(self):
        "Returns the field coordinates (gpu_process.py:851-857)."
        if self.piv_fields is None:
            self.init_piv_fields()
        x, y = self.piv_fields[-1].coords
        return _host(x * self.scaling_par), _host(y * self.scaling_par)

    @property
    def field_mask(self):
        """Returns the field mask for the velocity field."""
        return self.piv_fields[-1].field_mask

    @property
    def s2n_ratio(self):
        """Signal-to-noise ratio of the last pass of the last call with that
        pass's ``s2n_method`` / ``s2n_size``, 0 inside the mask
        (gpu_process.py:864-867, 1312-1376; unlike the reference it does not
        turn into ``None`` on a second access)."""
        if self._plan is None:
            return None
        m, n = self.piv_fields[-1].field_shape
        return self._plan.last_s2n(np.empty((m, n), self.dtype_f))

    def smoothing_log10_s(self, k):
        """log10 of the smoothing parameter chosen after pass ``k`` of the last call."""
        return self._plan.last_smoothing(k)

    def outlier_counts(self, k):
        """(before, after) replacement outlier counts of pass ``k`` of the last call."""
        return self._plan.last_counts(k)


def _flags_to_bool(a):
    if a is None:
        return None
    if isinstance(a, _native.DeviceBuffer):
        a = a.get()
    return np.asarray(a).astype(bool)


class piv_gpu:
    """Wrapper-class for PIVGPU that further applies input validation and
    provides the user interface (gpu_process.py:41-352).

    Parameters
    ----------
    frame_shape : tuple
        Shape of the images in pixels.
    min_search_size : int
        Width of the square search window in pixels.
    **kwargs
        PIV settings. See PIVGPU.

    Attributes
    ----------
    field_shape, coords, field_mask, val_locations, outliers
    """

    def __init__(self, frame_shape, min_search_size, **kwargs):
        unknown = set(kwargs) - set(_KWARGS) - {"device", "use_graph", "predictor", "max_reject_fraction",
                                                "predictor_pass"}
        if unknown:
            raise TypeError(f"unexpected keyword arguments: {sorted(unknown)}")
        cfg = dict(_DEFAULTS)
        cfg.update({k: v for k, v in kwargs.items() if k in _DEFAULTS})
        dtype_f = kwargs.get("dtype_f", "float32")
        _validate_settings(frame_shape, min_search_size, cfg)
        self.min_search_size = min_search_size
        self.frame_shape = frame_shape
        self.mask = cfg["mask"]
        self.frame_mask = (self.mask.astype(bool) if self.mask is not None
                           else np.full(self.frame_shape, fill_value=False, dtype=bool))
        self.dtype_f = DTYPE_f if dtype_f == "float64" else np.float32
        self.gpu_process = PIVGPU(frame_shape, min_search_size, **kwargs)

    def __call__(self, frame_a, frame_b):
        """Computes velocity field from an image pair; returns NumPy ``(u, v)``."""
        frames = [frame_a, frame_b]
        assert all(isinstance(f, np.ndarray) for f in frames) and \
            all(f.shape == self.frame_shape for f in frames) and \
            all(np.issubdtype(f.dtype, np.number) for f in frames) and \
            all(not np.iscomplex(f).any() for f in frames), \
            "Both frames must be an ndarray of {} values with shape {}.".format("real", self.frame_shape)
        return self.gpu_process.run_host(frame_a, frame_b)

    @property
    def field_shape(self):
        "Returns the field shape."
        return self.gpu_process.piv_fields[-1].field_shape

    @property
    def coords(self):
        "Returns the field coordinates."
        x, y = self.gpu_process.coords
        return x.get(), y.get()

    @property
    def field_mask(self):
        """Returns the field mask for the velocity field."""
        x, y = self.coords
        return self.frame_mask[y.astype(int), x.astype(int)]

    @property
    def val_locations(self):
        """Locations of the spurious vectors for the last iteration (before replacement)."""
        vl = self.gpu_process.init_val_locations
        if vl is not None:
            return _flags_to_bool(vl)
        return np.full(self.field_shape, fill_value=False, dtype=bool)

    @property
    def outliers(self):
        """Locations of the outliers after the last iteration."""
        vl = self.gpu_process.val_locations
        if vl is not None:
            return _flags_to_bool(vl)
        return np.full(self.field_shape, fill_value=False, dtype=bool)


gpu_piv = piv_gpu


def _validate_settings(frame_shape, min_search_size, c):
    """The argument checks of piv_gpu.__init__ (gpu_process.py:96-286), with
    the same AssertionError messages."""
    def tup(v, scalar_test, n):
        return (v,) * n if scalar_test(v) else v

    assert isinstance(min_search_size, int) and min_search_size >= 8, \
        "{} must be an {} number greater than 8.".format("min_search_size", "int")
    ssi = (c["search_size_iters"],) if isinstance(c["search_size_iters"], int) else c["search_size_iters"]
    assert isinstance(ssi, tuple) and len(ssi) > 0 and all(isinstance(i, int) for i in ssi) and \
        all(i > 0 for i in ssi), "{} must be a tuple of positive {} numbers.".format("search_size_iters", "int")
    n = len(ssi)
    max_ss = 2 ** (n - 1) * min_search_size
    assert isinstance(frame_shape, tuple) and len(frame_shape) == 2 and \
        all(isinstance(i, int) for i in frame_shape) and all(i >= max_ss for i in frame_shape), \
        "{} must be a tuple of {} values greater than {} of {}.".format("frame_shape", "int", "max_search_size",
                                                                       max_ss)
    ov = tup(c["overlap_ratio"], lambda v: v == 0 or isinstance(v, float), n)
    assert isinstance(ov, tuple) and len(ov) >= n and all(i == 0 or isinstance(i, float) for i in ov) and \
        all(0 <= i < 1 for i in ov), \
        "{} must be a tuple of {} values between 0 and 1, defined for all passes.".format("overlap_ratio", "real")
    sr = c["shrink_ratio"]
    assert sr == 1 or isinstance(sr, float) and 0 < sr <= 1, \
        "{} must be a {} number between 0 and 1.".format("shrink_ratio", "real")
    assert isinstance(c["center"], bool), "{} must have a {} value.".format("center", "bool")
    nz = tup(c["normalize"], lambda v: isinstance(v, bool), n)
    assert isinstance(nz, tuple) and len(nz) >= n and all(isinstance(i, bool) for i in nz), \
        "{} must be a tuple of {} values, defined for all passes.".format("normalize", "bool")
    spm = tup(c["subpixel_method"], lambda v: isinstance(v, str), n)
    assert isinstance(spm, tuple) and len(spm) >= n and all(i in ALLOWED_SUBPIXEL_METHODS for i in spm), \
        "{} for every pass must be one of {}.".format("subpixel_method", ALLOWED_SUBPIXEL_METHODS)
    nf = tup(c["n_fft"], lambda v: isinstance(v, int), n)
    assert isinstance(nf, tuple) and len(nf) >= n and all(isinstance(i, (int, float)) for i in nf) and \
        all(i >= 1 for i in nf), \
        "{} must be a tuple of {} numbers greater than 1, defined for all passes.".format("n_fft", "real")
    dp = tup(c["deforming_par"], lambda v: v == 0 or v == 1 or isinstance(v, float), n)
    assert isinstance(dp, tuple) and len(dp) >= n and all(i == 0 or i == 1 or isinstance(i, float) for i in dp) and \
        all(0 <= i <= 1 for i in dp), \
        "{} must be a tuple of {} values between 0 and 1, defined for all passes.".format("deforming_par", "real")
    bs = tup(c["batch_size"], lambda v: isinstance(v, int) or v == "full", n)
    assert isinstance(bs, tuple) and len(bs) >= n and all(isinstance(i, int) or i == "full" for i in bs), \
        "{} must be a tuple of positive {} numbers or 'full', defined for all passes.".format("batch_size", "int")
    sm = tup(c["s2n_method"], lambda v: isinstance(v, str), n)
    assert isinstance(sm, tuple) and len(sm) >= n and all(i in ALLOWED_S2N_METHODS for i in sm), \
        "{} for every pass must be one of {}.".format("s2n_method", ALLOWED_S2N_METHODS)
    for name in ("s2n_size", "validation_size"):
        t = tup(c[name], lambda v: isinstance(v, int), n)
        assert isinstance(t, tuple) and len(t) >= n and all(isinstance(i, int) for i in t) and \
            all(i > 0 for i in t), \
            "{} must be a tuple of positive {} numbers, defined for all passes.".format(name, "int")
    for name in ("s2n_tol", "median_tol", "mad_tol", "mean_tol", "rms_tol"):
        t = tup(c[name], lambda v: v is None or _is_real(v), n)
        assert isinstance(t, tuple) and len(t) >= n and all(i is None or _is_real(i) for i in t), \
            "{} must be a tuple of {} or None values, defined for all passes.".format(name, "real")
    nri = tup(c["num_replacing_iters"], lambda v: isinstance(v, int), n)
    assert isinstance(nri, tuple) and len(nri) >= n and all(isinstance(i, int) for i in nri) and \
        all(i >= 0 for i in nri), \
        "{} must be a tuple of positive {} numbers, defined for all passes.".format("num_replacing_iters", "int")
    rm = tup(c["replacing_method"], lambda v: isinstance(v, str), n)
    assert isinstance(rm, tuple) and len(rm) >= n and all(i in ALLOWED_REPLACING_METHODS for i in rm), \
        "{} for every pass must be one of {}.".format("replacing_method", ALLOWED_REPLACING_METHODS)
    rs = tup(c["replacing_size"], lambda v: isinstance(v, int), n)
    assert isinstance(rs, tuple) and len(rs) >= n and all(isinstance(i, int) for i in rs) and \
        all(i > 0 for i in rs), \
        "{} must be a tuple of positive {} values, defined for all passes.".format("replacing_size", "int")
    for name in ("revalidate", "smooth"):
        t = tup(c[name], lambda v: isinstance(v, bool), n)
        assert isinstance(t, tuple) and len(t) >= n and all(isinstance(i, bool) for i in t), \
            "{} must be a tuple of {} values, defined for all passes.".format(name, "bool")
    sp = tup(c["smoothing_par"], lambda v: v is None or _is_real(v), n)
    assert isinstance(sp, tuple) and len(sp) >= n and all(i is None or _is_real(i) for i in sp) and \
        all(i > 0 if i is not None else True for i in sp), \
        "{} must be a tuple of positive {} or None values, defined for all passes.".format("smoothing_par", "real")
    assert isinstance(c["dt"], int) or isinstance(c["dt"], float) and c["dt"] > 0, \
        "{} must be a {} number greater than 0.".format("dt", "real")
    assert isinstance(c["scaling_par"], int) or isinstance(c["scaling_par"], float) and c["scaling_par"] > 0, \
        "{} must be a {} number greater than 0.".format("scaling_par", "real")
    mask = c["mask"]
    assert mask is None or (isinstance(mask, np.ndarray) and mask.shape == frame_shape and
                            (np.issubdtype(mask.dtype, np.number) or mask.dtype == bool)), \
        "{} must be an ndarray of {} values with shape {}.".format("mask", "real", frame_shape)
