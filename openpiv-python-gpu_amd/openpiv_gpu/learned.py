"""Learned coarse-displacement predictor (no counterpart in the reference,
whose deformation predictor is always the previous pass's field,
gpu_process.py:620-625).

:class:`CoarseFlowNet` wraps the library's network (``piv_cnn_*``): a
4-level encoder-decoder CNN that maps a frame pair, box-averaged 8 x 8 and
standardised, to the displacement field on frame nodes (16 j + 8, 16 i + 8).
Training and inference run on the library's own HIP kernels (BF16 weights
and activations, FP32 accumulation, Adam).  Attached to a ``PIVGPU`` with
``predictor=``, it replaces the coarse passes: the last pass runs alone with
the network's field as its deformation predictor, and any pair whose median
test rejects more than ``max_reject_fraction`` of that pass's vectors runs the
full plain schedule instead.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _native

FACTOR = 8
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


class CoarseFlowNet:
    """The network for frames of ``frame_shape`` (multiples of 64 px)."""

    def __init__(self, frame_shape, max_batch: int = 1, params=None, seed: int = 0):
        ht, wd = (int(t) for t in frame_shape)
        if ht % (FACTOR * 8) or wd % (FACTOR * 8):
            raise ValueError("frame sides must be multiples of 64 px for the learned predictor")
        self.frame_shape = (ht, wd)
        self.in_shape = (ht // FACTOR, wd // FACTOR)
        self.out_shape = (self.in_shape[0] // 2, self.in_shape[1] // 2)
        self.max_batch = int(max_batch)
        self._lib = _native.load_library()
        h = ctypes.c_void_p()
        _native.check(self._lib.piv_cnn_create(self.in_shape[0], self.in_shape[1], self.max_batch, ctypes.byref(h)))
        self.handle = h
        n_layers = ctypes.c_int32()
        n_params = ctypes.c_int64()
        _native.check(self._lib.piv_cnn_layout(h, ctypes.byref(n_layers), None, None, ctypes.byref(n_params)))
        self.n_layers = n_layers.value
        self.n_params = n_params.value
        desc = np.zeros((self.n_layers, 8), np.int32)
        offs = np.zeros((self.n_layers, 2), np.int64)
        _native.check(self._lib.piv_cnn_layout(h, None, desc.ctypes.data, offs.ctypes.data, None))
        self.layer_desc = desc
        self.layer_offsets = offs
        self.set_params(self.he_init(seed) if params is None else params)

    def he_init(self, seed: int = 0):
        """He-normal weights (fan-in 9 * c_in), zero biases."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.n_params, np.float32)
        for (s0, s1, up, c0, c1, co, stride, relu), (w_off, b_off) in zip(self.layer_desc, self.layer_offsets):
            fan = 9 * (c0 + c1)
            p[w_off:w_off + co * fan] = rng.standard_normal(co * fan) * np.sqrt(2.0 / fan)
        return p

    # parameters ---------------------------------------------------------------
    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        _native.check(self._lib.piv_cnn_set_params(self.handle, p.ctypes.data))

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        _native.check(self._lib.piv_cnn_get_params(self.handle, p.ctypes.data))
        return p

    def get_grads(self):
        g = np.empty(self.n_params, np.float32)
        _native.check(self._lib.piv_cnn_get_grads(self.handle, g.ctypes.data))
        return g

    def get_buffer(self, buffer: int, batch: int, gradient: bool = False):
        """Activation (or gradient) of buffer 1..n_layers, as float32."""
        d = self.layer_desc[buffer - 1]
        h, w = self.buffer_shape(buffer)
        shape = (batch, 2, h, w) if (buffer == self.n_layers and not gradient) else (batch, h, w, int(d[5]))
        out = np.empty(shape, np.float32)
        _native.check(self._lib.piv_cnn_get_buffer(self.handle, buffer, int(gradient), batch, out.ctypes.data))
        return out

    def buffer_shape(self, buffer: int):
        h, w = self.in_shape
        shapes = [(h, w)]
        for (s0, s1, up, c0, c1, co, stride, relu) in self.layer_desc:
            hh, ww = shapes[s0]
            if up:
                hh, ww = 2 * hh, 2 * ww
            shapes.append((hh // stride, ww // stride))
        return shapes[buffer]

    def save(self, path):
        np.savez(path, params=self.get_params(), frame_shape=np.array(self.frame_shape),
                 layer_desc=self.layer_desc)

    @classmethod
    def load(cls, path, max_batch: int = 1):
        with np.load(path, allow_pickle=False) as z:
            net = cls(tuple(int(t) for t in z["frame_shape"]), max_batch=max_batch, params=z["params"])
            if not np.array_equal(z["layer_desc"], net.layer_desc):
                raise ValueError(f"{path}: saved architecture differs from the library's")
        return net

    # compute --------------------------------------------------------------------
    def prepare_input(self, frame_a, frame_b, out_ptr: int, stream=None):
        """Network input of a device-resident pair into ``out_ptr`` (BF16
        [H/8][W/8][2])."""
        _native.check(self._lib.piv_cnn_prepare_input(
            self.handle, frame_a.value, frame_b.value, _native.in_dtype_code(frame_a.dtype), *self.frame_shape,
            ctypes.c_void_p(out_ptr), None if stream is None else stream.handle))

    def forward(self, input_ptr: int, batch: int, out_ptr: int = None, stream=None):
        _native.check(self._lib.piv_cnn_forward(self.handle, ctypes.c_void_p(input_ptr), batch,
                                                None if out_ptr is None else ctypes.c_void_p(out_ptr),
                                                None if stream is None else stream.handle))

    def gradients(self, input_ptr: int, target_ptr: int, weight_ptr, batch: int, loss_ptr=None, stream=None):
        _native.check(self._lib.piv_cnn_gradients(
            self.handle, ctypes.c_void_p(input_ptr), ctypes.c_void_p(target_ptr),
            None if weight_ptr is None else ctypes.c_void_p(weight_ptr), batch,
            None if loss_ptr is None else ctypes.c_void_p(loss_ptr), None if stream is None else stream.handle))

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, stream=None):
        _native.check(self._lib.piv_cnn_adam_step(self.handle, lr, beta1, beta2, eps,
                                                  None if stream is None else stream.handle))

    def __del__(self):
        try:
            if getattr(self, "handle", None) is not None and self.handle.value:
                self._lib.piv_cnn_destroy(self.handle)
                self.handle = ctypes.c_void_p()
        except Exception:
            pass


def default_weights_path(frame_shape=(2048, 2048)):
    return os.path.join(WEIGHTS_DIR, f"coarse_flow_{frame_shape[0]}x{frame_shape[1]}.npz")
