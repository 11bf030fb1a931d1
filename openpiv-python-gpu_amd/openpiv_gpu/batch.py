"""Batches of image pairs with their own frame shapes and pass schedules.

No counterpart in the reference (one pair per ``PIVGPU.__call__``,
gpu_process.py:564-647).  Each pair is described by the keyword arguments a
``PIVGPU`` would take; the batch runs every pass level of all pairs as one
deformation launch and one correlation launch per FFT size, then each pair's
validation / replacement / smoothing, with the whole loop captured in one
hipGraph.  ``tune()`` picks launch shapes per level by timing them, and
keeps the choices in a tuning cache keyed by the GPU architecture and the
batch's description, so a later batch of the same description applies them
without timing.
"""

from __future__ import annotations

import ctypes
import hashlib
import json
import os

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
        self._key_src = {"pairs": [[list(map(int, shape)), int(mss), _jsonable(kw)] for shape, mss, kw in pairs],
                         "in_dtype": np.dtype(in_dtype).name}
        self._device = device
        self.tuning_source = None

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

    def choice_count(self) -> int:
        c = ctypes.c_int32()
        _native.check(self._lib.piv_batch_choice_count(self.handle, ctypes.byref(c)))
        return c.value

    def set_choices(self, choices):
        arr = np.ascontiguousarray(choices, np.int32)
        _native.check(self._lib.piv_batch_set_choices(self.handle, arr.ctypes.data, int(arr.size)))

    def tuning_key(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _native.check(self._lib.piv_device_arch(self._device, buf, 256))
        # layout 2: per level the deformation width, per group XCD flag and workgroup width
        src = dict(self._key_src, arch=buf.value.decode(), layout=2)
        return hashlib.sha1(json.dumps(src, sort_keys=True).encode()).hexdigest()

    def tune(self, repeats: int = 3, cache=True):
        """Launch shapes per level: from the tuning cache when it holds this
        batch's key, else timed (piv_batch_tune) and stored.  ``cache``:
        True for the default file (``$OPENPIV_HIP_TUNE_CACHE``, else
        ``~/.cache/openpiv_hip/batch_tuning.json``), a path, or False.
        Returns the choices (per level: deformation block width, then per
        correlation group its XCD-banded flag and workgroup width in waves);
        ``tuning_source`` says "cache" or "timed"."""
        path = _cache_path(cache)
        key = self.tuning_key() if path else None
        n = self.choice_count()
        if path:
            entry = _read_cache(path).get(key)
            if entry is not None and len(entry) == n:
                self.set_choices(entry)
                self.tuning_source = "cache"
                return np.asarray(entry, np.int32)
        out = np.zeros(max(n, 1), np.int32)
        _native.check(self._lib.piv_batch_tune(self.handle, repeats, out.ctypes.data))
        out = out[:n]
        self.tuning_source = "timed"
        if path:
            data = _read_cache(path)
            data[key] = [int(c) for c in out]
            _write_cache(path, data)
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


def _jsonable(kw):
    out = {}
    for k, v in sorted(dict(kw).items()):
        if isinstance(v, np.ndarray):
            v = hashlib.sha1(np.ascontiguousarray(v).tobytes()).hexdigest()
        elif isinstance(v, (tuple, list)):
            v = [x if isinstance(x, (int, float, str, bool)) or x is None else repr(x) for x in v]
        elif not (isinstance(v, (int, float, str, bool)) or v is None):
            v = repr(v)
        out[k] = v
    return out


def _cache_path(cache):
    if cache is False or cache is None:
        return None
    if cache is True:
        env = os.environ.get("OPENPIV_HIP_TUNE_CACHE")
        if env:
            return env
        return os.path.join(os.path.expanduser("~"), ".cache", "openpiv_hip", "batch_tuning.json")
    return os.fspath(cache)


def _read_cache(path):
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _write_cache(path, data):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


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
