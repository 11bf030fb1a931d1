"""Frame input for streams of image pairs.

The reference's notebooks read frames with OpenCV ``imread(path,
IMREAD_ANYDEPTH)`` and hand NumPy arrays to ``piv_gpu``, which uploads them
synchronously (``cp.asarray``, gpu_process.py:582-583).  Here frames are
decoded by the library (``read_tiff``: baseline uncompressed grayscale TIFF)
straight into pinned host buffers, uploaded on a copy stream while the
previous pair computes, and the plan waits for the upload's event, so file
decoding, PCIe transfer and GPU work of consecutive pairs overlap.
"""

from __future__ import annotations

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _native

_DTYPES = {8: np.uint8, 16: np.uint16, 32: np.float32}


def tiff_info(path):
    """``(height, width, dtype)`` of a baseline grayscale TIFF."""
    lib = _native.load_library()
    h, w, b = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    _native.check(lib.piv_tiff_info(os.fsencode(path), ctypes.byref(h), ctypes.byref(w), ctypes.byref(b)))
    return h.value, w.value, np.dtype(_DTYPES[b.value])


def read_tiff(path, out=None):
    """Decodes a baseline uncompressed grayscale TIFF (8/16-bit unsigned,
    32-bit float; strips; either byte order) into ``out`` or a new array."""
    h, w, dt = tiff_info(path)
    if out is None:
        out = np.empty((h, w), dt)
    if out.shape != (h, w) or out.dtype != dt or not out.flags.c_contiguous:
        raise ValueError(f"{path}: needs a contiguous {dt} array of shape {(h, w)}")
    _native.check(_native.load_library().piv_tiff_read(os.fsencode(path), out.ctypes.data, out.nbytes))
    return out


class PinnedArray:
    """A NumPy view of pinned (page-locked) host memory."""

    def __init__(self, shape, dtype):
        self._lib = _native.load_library()
        self.nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = ctypes.c_void_p()
        _native.check(self._lib.piv_host_alloc(ctypes.byref(p), self.nbytes))
        self.ptr = p
        buf = (ctypes.c_char * self.nbytes).from_address(p.value)
        self.array = np.frombuffer(buf, dtype=dtype).reshape(shape)

    def __del__(self):
        try:
            self.array = None
            self._lib.piv_host_free(self.ptr)
        except Exception:
            pass


class PairStream:
    """Runs a ``PIVGPU`` over a sequence of ``(path_a, path_b)`` TIFF pairs
    with overlapped decode / upload / compute.  Yields ``(index, u, v,
    flags)`` host arrays (the last pass, flags before replacement)."""

    def __init__(self, piv, depth: int = 3, workers: int = 4):
        self.piv = piv
        self.depth = max(2, depth)
        self.workers = workers
        self._lib = _native.load_library()
        self.copy_stream = _native.Stream()

    def run(self, pairs):
        pairs = list(pairs)
        if not pairs:
            return
        h, w, dt = tiff_info(pairs[0][0])
        if (h, w) != tuple(self.piv.frame_shape):
            raise ValueError(f"frames are {(h, w)}, the PIV object expects {self.piv.frame_shape}")
        plan = self.piv._get_plan(_native.in_dtype_code(dt))
        m, n = self.piv.piv_fields[-1].field_shape
        D = self.depth
        host = [(PinnedArray((h, w), dt), PinnedArray((h, w), dt)) for _ in range(D)]
        dev = [(_native.DeviceBuffer((h, w), dt), _native.DeviceBuffer((h, w), dt)) for _ in range(D)]
        events = [_native.Event() for _ in range(D)]
        nbytes = h * w * np.dtype(dt).itemsize

        def decode(i):
            s = i % D
            read_tiff(pairs[i][0], host[s][0].array)
            read_tiff(pairs[i][1], host[s][1].array)
            return i

        def upload(i):
            s = i % D
            for k in range(2):
                _native.check(self._lib.piv_memcpy_async(dev[s][k].ptr, host[s][k].ptr, nbytes, 0,
                                                         self.copy_stream.handle))
            events[s].record(self.copy_stream)

        with ThreadPoolExecutor(self.workers) as ex:
            futures = {i: ex.submit(decode, i) for i in range(min(D - 1, len(pairs)))}
            futures[0].result()
            upload(0)
            for i in range(len(pairs)):
                s = i % D
                if i + 1 < len(pairs):
                    futures[i + 1].result()
                    upload(i + 1)                       # overlaps pair i's compute
                if i + D - 1 < len(pairs):
                    # slot (i + D - 1) % D was last uploaded for pair i - 1, whose
                    # run has returned, so its host buffers are free
                    futures[i + D - 1] = ex.submit(decode, i + D - 1)
                u = np.empty((m, n), np.float32)
                v = np.empty((m, n), np.float32)
                vi = np.empty((m, n), np.uint8)
                _native.check(self._lib.piv_plan_wait_event(plan.handle, events[s].handle))
                plan.run(dev[s][0].value, dev[s][1].value, True, u.ctypes.data, v.ctypes.data, vi.ctypes.data,
                         None, False)
                futures.pop(i, None)
                yield i, u, v, vi.astype(bool)
