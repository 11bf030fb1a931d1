"""Batches of image pairs with their own frame shapes and pass schedules.

No counterpart in the reference (one pair per ``PIVGPU.__call__``,
gpu_process.py:564-647).  Each pair is described by the keyword arguments a
``PIVGPU`` would take; the batch runs every pass level of all pairs as one
deformation launch and one correlation launch per FFT size, then each pair's
validation / replacement / smoothing, with the whole loop captured in one
hipGraph.  ``tune()`` picks launch shapes per level by timing them.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .gpu_process import PIVGPU


class PIVBatch:
    def __init__(self, pairs, in_dtype=np.uint16, device: int = 0):
        """``pairs``: sequence of ``(frame_shape, min_search_size, kwargs)``."""
        self._lib = _native.load_library()
        _native.require_device()
        self.pivs = [PIVGPU(tuple(shape), mss, dtype_f="float32", **dict(kw)) for shape, mss, kw in pairs]
        code = _native.in_dtype_code(np.dtype(in_dtype))
        descs = (_native.PlanDesc * len(self.pivs))(*[p._plan_desc(code) for p in self.pivs])
        h = ctypes.c_void_p()
        _native.check(self._lib.piv_batch_create(device, len(self.pivs), descs, ctypes.byref(h)))
        self.handle = h
        self.n = len(self.pivs)
        self._frames = None

    def set_frames(self, frames):
        """``frames``: sequence of ``(frame_a, frame_b)`` device buffers (kept
        alive by the batch and read in place on every run)."""
        if len(frames) != self.n:
            raise ValueError("one (frame_a, frame_b) per pair")
        self._frames = list(frames)
        arr_a = (ctypes.c_void_p * self.n)(*[f[0].value for f in frames])
        arr_b = (ctypes.c_void_p * self.n)(*[f[1].value for f in frames])
        _native.check(self._lib.piv_batch_set_frames(self.handle, arr_a, arr_b))

    def run(self):
        _native.check(self._lib.piv_batch_run(self.handle))

    def tune(self, repeats: int = 3):
        """Times the launch-shape candidates; returns the choices per level."""
        out = np.zeros(64, np.int32)
        _native.check(self._lib.piv_batch_tune(self.handle, repeats, out.ctypes.data))
        return out

    def field_shape(self, i):
        m, n = ctypes.c_int32(), ctypes.c_int32()
        _native.check(self._lib.piv_batch_pair_shape(self.handle, i, ctypes.byref(m), ctypes.byref(n)))
        return m.value, n.value

    def result(self, i):
        """Host ``(u, v, val_init, val_final)`` of pair ``i`` (last pass)."""
        m, n = self.field_shape(i)
        u = np.empty((m, n), np.float32)
        v = np.empty((m, n), np.float32)
        vi = np.empty((m, n), np.uint8)
        vf = np.empty((m, n), np.uint8)
        _native.check(self._lib.piv_batch_get(self.handle, i, u.ctypes.data, v.ctypes.data, vi.ctypes.data,
                                              vf.ctypes.data, 0))
        return u, v, vi.astype(bool), vf.astype(bool)

    def __del__(self):
        try:
            if getattr(self, "handle", None) is not None and self.handle.value:
                self._lib.piv_batch_destroy(self.handle)
                self.handle = ctypes.c_void_p()
        except Exception:
            pass


def uneven_pair_spec(seed: int, sizes=((512, 512), (768, 1024), (2048, 2048), (2048, 3072), (4096, 4096))):
    """The uneven-batch workload's draw for one seed: frame shape (rows,
    cols), schedule 64->32->16 or 48->24, overlap ~ U(0.5, 0.75)."""
    rng = np.random.Generator(np.random.PCG64([seed, 7]))
    shape = sizes[int(rng.integers(len(sizes)))]
    if rng.random() < 0.5:
        mss, iters = 16, (1, 1, 1)
    else:
        mss, iters = 24, (1, 1)
    overlap = float(rng.uniform(0.5, 0.75))
    return shape, mss, dict(search_size_iters=iters, overlap_ratio=overlap, median_tol=2)
