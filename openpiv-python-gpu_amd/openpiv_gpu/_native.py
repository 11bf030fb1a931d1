"""ctypes binding of ``libopenpiv_hip.so`` (declared in include/openpiv_hip.h).

There is no CPU fallback: importing a compute entry point without the built
library, or calling one without a visible GPU, raises.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

MAX_PASSES = 16
PIV_U8, PIV_U16, PIV_F32, PIV_F64 = 0, 1, 2, 3
PIV_FP32, PIV_FP64 = 0, 1
SUBPIXEL_CODES = {"gaussian": 0, "parabolic": 1, "centroid": 2}
S2N_CODES = {"peak2peak": 0, "peak2mean": 1, "peak2energy": 2}
REPLACING_CODES = {"spring": 0, "median": 1, "mean": 2}
MAX_REPLACING_ITERS = 16

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PIV_HIP_LIB", os.path.join(_HERE, "_lib", "libopenpiv_hip.so"))

# Every symbol include/openpiv_hip.h declares.
EXPORTED_SYMBOLS = (
    "piv_last_error", "piv_device_count", "piv_plan_create", "piv_plan_destroy", "piv_run", "piv_run_async",
    "piv_plan_synchronize",
    "piv_plan_last_smoothing", "piv_plan_last_counts", "piv_stage_correlate", "piv_stage_deform",
    "piv_stage_validate", "piv_stage_replace",
    "piv_stage_grid_interp", "piv_stage_median_validate", "piv_stage_spring_replace", "piv_stage_smoothn",
    "piv_set_device", "piv_malloc", "piv_free", "piv_memcpy", "piv_memset", "piv_synchronize",
    "piv_stream_create", "piv_stream_destroy", "piv_stream_synchronize", "piv_event_create", "piv_event_destroy",
    "piv_event_record", "piv_event_elapsed_ms", "piv_stage_correlate_dft", "piv_synth_pair",
    "piv_cnn_create", "piv_cnn_destroy", "piv_cnn_layout", "piv_cnn_set_params", "piv_cnn_get_params",
    "piv_cnn_get_grads", "piv_cnn_get_buffer", "piv_cnn_forward", "piv_cnn_gradients", "piv_cnn_adam_step",
    "piv_cnn_prepare_input", "piv_plan_set_predictor", "piv_plan_last_fallback",
    "piv_batch_create", "piv_batch_destroy", "piv_batch_set_frames", "piv_batch_run", "piv_batch_tune",
    "piv_batch_pair_shape", "piv_batch_get", "piv_tiff_info", "piv_tiff_read", "piv_host_alloc", "piv_host_free",
    "piv_memcpy_async", "piv_stream_wait_event", "piv_plan_wait_event", "piv_plan_set_mask",
    "piv_plan_last_smoothing_iterations", "piv_stage_smoothn_masked", "piv_stage_correlate64",
    "piv_plan_last_s2n", "piv_stage_replace_noalloc", "piv_batch_choice_count", "piv_batch_set_choices",
    "piv_device_arch", "piv_plan_learned_counts",
)

# largest FFT size of the correlation kernels (planes above 128 points live in
# device-memory slabs, 512 MiB per plan)
MAX_FFT_SIZE = 1024
# largest validation_size / replacing_size the kernels take: (2 * 8 + 1)^2 - 1
# = 288 neighbours, 64 spring residue classes
MAX_NEIGHBOURHOOD_SIZE = 8
SYNTH_MAX_VORTICES = 8
FIELD_UNIFORM, FIELD_SINE, FIELD_VORTICES = 0, 1, 2


class PassDesc(ctypes.Structure):
    _fields_ = [
        ("search_size", ctypes.c_int32),
        ("spacing", ctypes.c_int32),
        ("x_offset", ctypes.c_int32),
        ("y_offset", ctypes.c_int32),
        ("field_m", ctypes.c_int32),
        ("field_n", ctypes.c_int32),
        ("fft_size", ctypes.c_int32),
        ("normalize", ctypes.c_int32),
        ("subpixel", ctypes.c_int32),
        ("deforming_par", ctypes.c_float),
        ("validate", ctypes.c_int32),
        ("median_tol", ctypes.c_float),
        ("num_replacing_iters", ctypes.c_int32),
        ("smooth", ctypes.c_int32),
        ("smoothing_par", ctypes.c_double),
        ("s2n_method", ctypes.c_int32),
        ("s2n_size", ctypes.c_int32),
        ("s2n_tol", ctypes.c_float),
        ("mad_tol", ctypes.c_float),
        ("mean_tol", ctypes.c_float),
        ("rms_tol", ctypes.c_float),
        ("replacing_method", ctypes.c_int32),
        ("revalidate", ctypes.c_int32),
        ("window_size", ctypes.c_int32),
        ("validation_size", ctypes.c_int32),
        ("replacing_size", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("in_dtype", ctypes.c_int32),
        ("n_passes", ctypes.c_int32),
        ("passes", PassDesc * MAX_PASSES),
        ("out_scale", ctypes.c_float),
        ("use_graph", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("scaling_par", ctypes.c_double),
        ("dt", ctypes.c_double),
    ]


class SynthField(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("n_vortices", ctypes.c_int32),
        ("u0", ctypes.c_double),
        ("v0", ctypes.c_double),
        ("amp_u", ctypes.c_double),
        ("amp_v", ctypes.c_double),
        ("period", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("cx", ctypes.c_double * SYNTH_MAX_VORTICES),
        ("cy", ctypes.c_double * SYNTH_MAX_VORTICES),
        ("rc", ctypes.c_double * SYNTH_MAX_VORTICES),
        ("circ", ctypes.c_double * SYNTH_MAX_VORTICES),
    ]


class SynthModel(ctypes.Structure):
    _fields_ = [
        ("sigma", ctypes.c_double),
        ("i0_low", ctypes.c_double),
        ("i0_high", ctypes.c_double),
        ("noise_sigma", ctypes.c_double),
        ("max_value", ctypes.c_double),
        ("out_dtype", ctypes.c_int32),
        ("margin", ctypes.c_int32),
        ("n_particles", ctypes.c_int64),
    ]


class NativeError(RuntimeError):
    pass


_lib: Optional[ctypes.CDLL] = None


def _bind(lib):
    vp, i32, f32, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
    sig = {
        "piv_last_error": (ctypes.c_char_p, []),
        "piv_device_count": (i32, []),
        "piv_plan_create": (i32, [i32, ctypes.POINTER(PlanDesc), ctypes.POINTER(vp)]),
        "piv_plan_destroy": (None, [vp]),
        "piv_run": (i32, [vp, vp, vp, i32, vp, vp, vp, vp, i32]),
        "piv_run_async": (i32, [vp, vp, vp, vp, vp, vp, vp]),
        "piv_plan_synchronize": (i32, [vp]),
        "piv_plan_learned_counts": (i32, [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
        "piv_plan_last_smoothing": (i32, [vp, i32, ctypes.POINTER(f64)]),
        "piv_plan_last_counts": (i32, [vp, i32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
        "piv_stage_correlate": (i32, [vp, vp, i32, i32, i32, ctypes.POINTER(PassDesc), vp, vp, vp, vp, vp, vp, vp,
                                      vp, vp]),
        "piv_stage_validate": (i32, [vp, vp, i32, i32, vp, ctypes.POINTER(PassDesc), vp, vp, vp, vp]),
        "piv_stage_replace": (i32, [vp, vp, vp, i32, i32, ctypes.POINTER(PassDesc), vp, vp]),
        "piv_stage_replace_noalloc": (i32, [vp, vp, vp, i32, i32, ctypes.POINTER(PassDesc), vp, vp]),
        "piv_stage_deform": (i32, [vp, vp, i32, i32, i32, ctypes.POINTER(PassDesc), vp, vp, vp, vp, vp]),
        "piv_stage_grid_interp": (i32, [ctypes.POINTER(PassDesc), ctypes.POINTER(PassDesc), vp, vp, vp, vp, vp]),
        "piv_stage_median_validate": (i32, [vp, vp, i32, i32, f32, vp, vp, vp]),
        "piv_stage_spring_replace": (i32, [vp, vp, vp, vp, i32, i32, vp]),
        "piv_stage_smoothn": (i32, [vp, vp, i32, i32, f64, vp, vp]),
        "piv_stage_smoothn_masked": (i32, [vp, vp, vp, i32, i32, f64, vp, vp, vp]),
        "piv_stage_correlate64": (i32, [vp, vp, i32, i32, i32, ctypes.POINTER(PassDesc), vp, vp, vp, vp, vp, vp, vp,
                                        vp, vp]),
        "piv_plan_set_mask": (i32, [vp, vp, i32, i32]),
        "piv_plan_last_s2n": (i32, [vp, vp]),
        "piv_plan_last_smoothing_iterations": (i32, [vp, i32, ctypes.POINTER(ctypes.c_int32)]),
        "piv_stage_correlate_dft": (i32, [vp, vp, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp]),
        "piv_synth_pair": (i32, [ctypes.POINTER(SynthField), ctypes.POINTER(SynthModel), ctypes.c_uint64, i32, i32,
                                 vp, vp, vp, vp, vp]),
        "piv_cnn_create": (i32, [i32, i32, i32, ctypes.POINTER(vp)]),
        "piv_cnn_destroy": (None, [vp]),
        "piv_cnn_layout": (i32, [vp, vp, vp, vp, vp]),
        "piv_cnn_set_params": (i32, [vp, vp]),
        "piv_cnn_get_params": (i32, [vp, vp]),
        "piv_cnn_get_grads": (i32, [vp, vp]),
        "piv_cnn_get_buffer": (i32, [vp, i32, i32, i32, vp]),
        "piv_cnn_forward": (i32, [vp, vp, i32, vp, vp]),
        "piv_cnn_gradients": (i32, [vp, vp, vp, vp, i32, vp, vp]),
        "piv_cnn_adam_step": (i32, [vp, f32, f32, f32, f32, vp]),
        "piv_cnn_prepare_input": (i32, [vp, vp, vp, i32, i32, i32, vp, vp]),
        "piv_plan_set_predictor": (i32, [vp, vp, i32, f32]),
        "piv_plan_last_fallback": (i32, [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
        "piv_batch_create": (i32, [i32, i32, ctypes.POINTER(PlanDesc), ctypes.POINTER(vp)]),
        "piv_batch_destroy": (None, [vp]),
        "piv_batch_set_frames": (i32, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "piv_batch_run": (i32, [vp]),
        "piv_batch_tune": (i32, [vp, i32, vp]),
        "piv_batch_choice_count": (i32, [vp, vp]),
        "piv_batch_set_choices": (i32, [vp, vp, i32]),
        "piv_device_arch": (i32, [i32, vp, i32]),
        "piv_batch_pair_shape": (i32, [vp, i32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
        "piv_batch_get": (i32, [vp, i32, vp, vp, vp, vp, i32]),
        "piv_tiff_info": (i32, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                ctypes.POINTER(ctypes.c_int32)]),
        "piv_tiff_read": (i32, [ctypes.c_char_p, vp, ctypes.c_size_t]),
        "piv_host_alloc": (i32, [ctypes.POINTER(vp), ctypes.c_size_t]),
        "piv_host_free": (i32, [vp]),
        "piv_memcpy_async": (i32, [vp, vp, ctypes.c_size_t, i32, vp]),
        "piv_stream_wait_event": (i32, [vp, vp]),
        "piv_plan_wait_event": (i32, [vp, vp]),
        "piv_set_device": (i32, [i32]),
        "piv_malloc": (i32, [ctypes.POINTER(vp), ctypes.c_size_t]),
        "piv_free": (i32, [vp]),
        "piv_memcpy": (i32, [vp, vp, ctypes.c_size_t, i32]),
        "piv_memset": (i32, [vp, i32, ctypes.c_size_t]),
        "piv_synchronize": (i32, []),
        "piv_stream_create": (i32, [ctypes.POINTER(vp)]),
        "piv_stream_destroy": (i32, [vp]),
        "piv_stream_synchronize": (i32, [vp]),
        "piv_event_create": (i32, [ctypes.POINTER(vp)]),
        "piv_event_destroy": (i32, [vp]),
        "piv_event_record": (i32, [vp, vp]),
        "piv_event_elapsed_ms": (i32, [vp, vp, ctypes.POINTER(ctypes.c_float)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load_library(path: str = None) -> ctypes.CDLL:
    """Loads (once) and returns the HIP library; raises if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"HIP library not built: {p} (run build_native.py or __graft_entry__.build())")
    lib = ctypes.CDLL(p)
    _bind(lib)
    if path is None:
        _lib = lib
    return lib


def check(code: int):
    if code != 0:
        msg = load_library().piv_last_error()
        raise NativeError(msg.decode() if msg else f"openpiv_hip error {code}")


def device_count() -> int:
    return int(load_library().piv_device_count())


def require_device():
    if device_count() <= 0:
        raise NativeError("no HIP device visible: the PIV hot path runs only on the GPU")


def in_dtype_code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.uint8:
        return PIV_U8
    if dt == np.uint16:
        return PIV_U16
    if dt == np.float32:
        return PIV_F32
    if dt == np.float64:
        return PIV_F64
    raise ValueError(f"unsupported frame dtype {dt}")


class DeviceBuffer:
    """Device allocation owned by Python (freed with the object)."""

    def __init__(self, shape, dtype):
        lib = load_library()
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib.piv_malloc(ctypes.byref(p), self.nbytes))
        self._lib = lib
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        buf = cls(a.shape, a.dtype)
        buf.upload(a)
        return buf

    @classmethod
    def zeros(cls, shape, dtype):
        buf = cls(shape, dtype)
        check(buf._lib.piv_memset(buf.ptr, 0, buf.nbytes))
        return buf

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(self._lib.piv_memcpy(self.ptr, a.ctypes.data, self.nbytes, 0))

    def get(self):
        out = np.empty(self.shape, self.dtype)
        check(self._lib.piv_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    @property
    def value(self):
        return self.ptr.value

    def view(self, index):
        """Sub-array ``index`` along the first axis (no copy)."""
        sub = self.shape[1:]
        step = int(np.prod(sub)) * self.dtype.itemsize
        return DeviceView(self.value + index * step, sub, self.dtype, self)

    def free(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            self._lib.piv_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView:
    """A typed window into device memory owned elsewhere (no ownership)."""

    def __init__(self, ptr: int, shape, dtype, owner=None):
        self.ptr = ctypes.c_void_p(int(ptr))
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._owner = owner
        self._lib = load_library()

    @property
    def value(self):
        return self.ptr.value

    def get(self):
        out = np.empty(self.shape, self.dtype)
        check(self._lib.piv_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    def view(self, index):
        """Sub-array ``index`` along the first axis."""
        sub = self.shape[1:]
        step = int(np.prod(sub)) * self.dtype.itemsize
        return DeviceView(self.value + index * step, sub, self.dtype, self._owner or self)


class Stream:
    def __init__(self):
        self._lib = load_library()
        h = ctypes.c_void_p()
        check(self._lib.piv_stream_create(ctypes.byref(h)))
        self.handle = h

    def synchronize(self):
        check(self._lib.piv_stream_synchronize(self.handle))

    def __del__(self):
        try:
            self._lib.piv_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        self._lib = load_library()
        h = ctypes.c_void_p()
        check(self._lib.piv_event_create(ctypes.byref(h)))
        self.handle = h

    def record(self, stream=None):
        check(self._lib.piv_event_record(self.handle, stream.handle if stream is not None else None))

    def elapsed_ms(self, stop):
        ms = ctypes.c_float()
        check(self._lib.piv_event_elapsed_ms(self.handle, stop.handle, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            self._lib.piv_event_destroy(self.handle)
        except Exception:
            pass


def synchronize():
    check(load_library().piv_synchronize())


class Plan:
    """Owns one native plan (frame shape + pass schedule)."""

    def __init__(self, desc: PlanDesc, device: int = 0):
        lib = load_library()
        require_device()
        handle = ctypes.c_void_p()
        check(lib.piv_plan_create(device, ctypes.byref(desc), ctypes.byref(handle)))
        self._lib = lib
        self.handle = handle
        self.desc = desc

    def run(self, frame_a, frame_b, frames_on_device, u, v, val_init, val_final, outputs_on_device):
        check(self._lib.piv_run(self.handle, frame_a, frame_b, int(frames_on_device), u, v, val_init, val_final,
                                int(outputs_on_device)))

    def run_async(self, frame_a, frame_b, u, v, val_init=None, val_final=None):
        """Queues one run on device frames and device outputs; see synchronize()."""
        check(self._lib.piv_run_async(self.handle, frame_a, frame_b, u, v, val_init, val_final))

    def synchronize(self):
        check(self._lib.piv_plan_synchronize(self.handle))

    def last_smoothing(self, k: int) -> float:
        out = ctypes.c_double()
        check(self._lib.piv_plan_last_smoothing(self.handle, k, ctypes.byref(out)))
        return out.value

    def last_counts(self, k: int):
        a, b = ctypes.c_int32(), ctypes.c_int32()
        check(self._lib.piv_plan_last_counts(self.handle, k, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_mask(self, mask):
        """Frame mask (bool/number array of the frame shape; ``None`` removes it)."""
        if mask is None:
            check(self._lib.piv_plan_set_mask(self.handle, None, 0, 0))
            return
        m = np.ascontiguousarray(np.asarray(mask).astype(bool).view(np.uint8))
        check(self._lib.piv_plan_set_mask(self.handle, m.ctypes.data, int(m.shape[0]), int(m.shape[1])))

    def last_s2n(self, out):
        """Signal-to-noise ratio of the last pass of the last run into ``out``
        (a C-contiguous host array of the field shape and plan precision)."""
        check(self._lib.piv_plan_last_s2n(self.handle, out.ctypes.data))
        return out

    def last_smoothing_iterations(self, k: int) -> int:
        out = ctypes.c_int32()
        check(self._lib.piv_plan_last_smoothing_iterations(self.handle, k, ctypes.byref(out)))
        return out.value

    def set_predictor(self, net, max_reject_fraction: float, first_pass: int = -1):
        """Attach (or with ``net=None`` detach) a learned predictor network
        feeding pass ``first_pass`` (-1: the last pass)."""
        self._net = net  # keep the network alive while the plan uses it
        check(self._lib.piv_plan_set_predictor(self.handle, None if net is None else net.handle, int(first_pass),
                                               float(max_reject_fraction)))

    def learned_counts(self):
        """(learned runs, fallbacks to the full schedule) of the plan so far."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(self._lib.piv_plan_learned_counts(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def last_fallback(self):
        """``(fell_back, rejected)`` of the last run with a predictor."""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        check(self._lib.piv_plan_last_fallback(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return bool(a.value), b.value

    def close(self):
        if getattr(self, "handle", None):
            self._lib.piv_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
