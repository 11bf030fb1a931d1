"""Seeded synthetic PIV image pairs.

The reference ships no generator and no tests; this module defines the
particle-image model the benchmarks and parity tests run on:

* particles are isotropic Gaussian blobs ``I0 * exp(-r^2 / (2 sigma^2))``
  point-sampled at integer pixel centres (pixel ``(r, c)`` sits at
  ``x = c, y = r``, the pixel grid the reference warps on,
  gpu_process.py:1153-1155), cut off at ``r <= 4 sigma``;
* a fixed particle count ``round(density * area)`` is drawn over the frame
  extended by a margin of ``ceil(max|d|) + ceil(4 sigma) + 1`` pixels;
* placement is symmetric: frame A holds each particle at ``p - d(p)/2``,
  frame B at ``p + d(p)/2``, so the displacement at a node is ``d(node)``;
* independent additive Gaussian noise per frame, then ``rint``, clip, cast.

All randomness comes from ``numpy.random.Generator(PCG64(seed))`` in a fixed
draw order (positions x, positions y, peak intensities, noise A, noise B).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional, Tuple

import numpy as np

DisplacementFn = Callable[[np.ndarray, np.ndarray], Tuple[np.ndarray, np.ndarray]]


def uniform_displacement(u: float, v: float) -> DisplacementFn:
    """Constant displacement (u along x / columns, v along y / rows)."""
    def fn(x, y):
        return np.full_like(x, u, dtype=np.float64), np.full_like(y, v, dtype=np.float64)
    fn.max_abs = max(abs(u), abs(v))
    fn.params = {"kind": "uniform", "u0": u, "v0": v}
    return fn


def sine_displacement(amp_u: float = 4.0, amp_v: float = 3.0, period: float = 256.0) -> DisplacementFn:
    """``u = amp_u sin(2 pi y / period)``, ``v = amp_v cos(2 pi x / period)``."""
    def fn(x, y):
        return (amp_u * np.sin(2.0 * np.pi * y / period),
                amp_v * np.cos(2.0 * np.pi * x / period))
    fn.max_abs = max(abs(amp_u), abs(amp_v))
    fn.params = {"kind": "sine", "amp_u": amp_u, "amp_v": amp_v, "period": period}
    return fn


def lamb_oseen_displacement(rng: np.random.Generator, shape: Tuple[int, int]) -> DisplacementFn:
    """Random superposition of 1-4 Lamb-Oseen vortices plus a uniform shear.

    Core radius ``r_c ~ U(20, 200)`` px, peak swirl ``~ U(2, 10)`` px (the
    swirl peaks at ``0.6382 * Gamma / (2 pi r_c)``), shear ``|gamma| <= 0.02``
    px/px applied as ``u += gamma (y - H/2)``.
    """
    ht, wd = shape
    n_vort = int(rng.integers(1, 5))
    cx = rng.uniform(0, wd, n_vort)
    cy = rng.uniform(0, ht, n_vort)
    rc = rng.uniform(20.0, 200.0, n_vort)
    peak = rng.uniform(2.0, 10.0, n_vort) * rng.choice([-1.0, 1.0], n_vort)
    circ = peak * 2.0 * np.pi * rc / 0.6382
    gamma = rng.uniform(-0.02, 0.02)

    def fn(x, y):
        u = gamma * (y - ht / 2.0)
        v = np.zeros_like(x, dtype=np.float64)
        for k in range(n_vort):
            dx, dy = x - cx[k], y - cy[k]
            r2 = dx * dx + dy * dy
            r2s = np.where(r2 > 0, r2, 1.0)
            ut = circ[k] / (2.0 * np.pi) * (1.0 - np.exp(-r2s / rc[k] ** 2)) / r2s
            ut = np.where(r2 > 0, ut, 0.0)
            u = u - ut * dy
            v = v + ut * dx
        return u, v
    fn.max_abs = float(np.max(np.abs(peak))) + 0.02 * max(ht, wd) / 2.0
    fn.params = {"kind": "vortices", "gamma": gamma, "cx": cx, "cy": cy, "rc": rc, "circ": circ}
    return fn


@dataclass
class ParticleModel:
    sigma: float = 0.65
    density: float = 0.05
    i0_low: float = 150.0
    i0_high: float = 255.0
    noise_sigma: float = 2.0
    max_value: float = 255.0
    dtype: type = np.uint8


def _render(ht, wd, xp, yp, i0, sigma):
    """Point-samples Gaussian blobs at integer pixel centres (r <= 4 sigma)."""
    rad = int(math.ceil(4.0 * sigma))
    offs = np.arange(-rad, rad + 1)
    oy, ox = np.meshgrid(offs, offs, indexing="ij")
    oy = oy.reshape(-1)
    ox = ox.reshape(-1)
    img = np.zeros(ht * wd, dtype=np.float64)
    cx = np.rint(xp).astype(np.int64)
    cy = np.rint(yp).astype(np.int64)
    # Chunk over particles to bound memory.
    chunk = 65536
    for s in range(0, xp.size, chunk):
        sl = slice(s, s + chunk)
        px = cx[sl, None] + ox[None, :]
        py = cy[sl, None] + oy[None, :]
        d2 = (px - xp[sl, None]) ** 2 + (py - yp[sl, None]) ** 2
        ok = (d2 <= (4.0 * sigma) ** 2) & (px >= 0) & (px < wd) & (py >= 0) & (py < ht)
        val = i0[sl, None] * np.exp(-d2 / (2.0 * sigma * sigma))
        np.add.at(img, (py[ok] * wd + px[ok]), val[ok])
    return img.reshape(ht, wd)


def make_pair(shape: Tuple[int, int], disp: DisplacementFn, model: ParticleModel,
              seed: int, rng: Optional[np.random.Generator] = None):
    """Returns ``(frame_a, frame_b)`` of ``model.dtype`` for one seeded pair."""
    ht, wd = shape
    rng = np.random.Generator(np.random.PCG64(seed)) if rng is None else rng
    margin = int(math.ceil(getattr(disp, "max_abs", 0.0))) + int(math.ceil(4.0 * model.sigma)) + 1
    area = (ht + 2 * margin) * (wd + 2 * margin)
    n = int(round(model.density * area))
    px = rng.uniform(-margin, wd + margin, n)
    py = rng.uniform(-margin, ht + margin, n)
    i0 = rng.uniform(model.i0_low, model.i0_high, n)
    du, dv = disp(px, py)
    fa = _render(ht, wd, px - du / 2.0, py - dv / 2.0, i0, model.sigma)
    fb = _render(ht, wd, px + du / 2.0, py + dv / 2.0, i0, model.sigma)
    out = []
    for f in (fa, fb):
        if model.noise_sigma > 0:
            f = f + rng.normal(0.0, model.noise_sigma, f.shape)
        f = np.clip(np.rint(f), 0, model.max_value)
        out.append(f.astype(model.dtype))
    return out[0], out[1]


# Named workloads -----------------------------------------------------------

def cfg_512_uint8(seed: int = 1):
    """512x512 uint8 pair, sigma 0.65, density 0.05, I0 ~ U(150,255), noise 2,
    uniform shift (u, v) = (2.35, -1.60). Returns ``(a, b, disp_fn)``."""
    disp = uniform_displacement(2.35, -1.60)
    a, b = make_pair((512, 512), disp, ParticleModel(), seed)
    return a, b, disp


def cfg_2048_uint16(seed: int = 2, shape=(2048, 2048)):
    """2048x2048 uint16 pair with 12-bit intensities (I0 ~ U(1000,4095),
    noise 20, clip 0..4095) and ``u = 4 sin(2 pi y/256)``,
    ``v = 3 cos(2 pi x/256)``. Returns ``(a, b, disp_fn)``."""
    disp = sine_displacement(4.0, 3.0, 256.0)
    model = ParticleModel(sigma=0.65, density=0.05, i0_low=1000.0, i0_high=4095.0,
                          noise_sigma=20.0, max_value=4095.0, dtype=np.uint16)
    a, b = make_pair(shape, disp, model, seed)
    return a, b, disp


def cfg_vortex_uint16(seed: int, shape: Tuple[int, int]):
    """Uneven-batch pair: Lamb-Oseen vortices + shear, 12-bit uint16 model."""
    rng = np.random.Generator(np.random.PCG64(seed))
    disp = lamb_oseen_displacement(rng, shape)
    model = ParticleModel(sigma=0.65, density=0.05, i0_low=1000.0, i0_high=4095.0,
                          noise_sigma=20.0, max_value=4095.0, dtype=np.uint16)
    a, b = make_pair(shape, disp, model, seed, rng=rng)
    return a, b, disp


def _window_chunk(ws, seed, c, k, wa, wb, shifts):
    """Renders windows ``c*chunk .. c*chunk+k`` of :func:`window_pairs` into
    ``wa``/``wb``/``shifts`` (views of that chunk) from the chunk's own stream."""
    rng = np.random.Generator(np.random.PCG64([seed, c]))
    shifts[:] = rng.uniform(-8.0, 8.0, (k, 2))
    sigmas = rng.uniform(0.5, 1.2, k)
    noise = rng.uniform(0.0, 0.05, k)
    # each frame moves a particle by at most 4 px and a blob reaches 4 sigma
    margin = 4 + 5
    npart = int(round(0.05 * (ws + 2 * margin) ** 2))
    px = rng.uniform(-margin, ws + margin, (k, npart))
    py = rng.uniform(-margin, ws + margin, (k, npart))
    i0 = rng.uniform(0.5, 1.0, (k, npart))
    grid = np.arange(ws, dtype=np.float64)
    s2 = 2.0 * sigmas[:, None, None] ** 2
    for sign, dst in ((-0.5, wa), (0.5, wb)):
        cx = px + sign * shifts[:, 0:1]
        cy = py + sign * shifts[:, 1:2]
        gy = np.exp(-(grid[None, :, None] - cy[:, None, :]) ** 2 / s2) * i0[:, None, :]   # (k, ws, P)
        gx = np.exp(-(grid[None, None, :] - cx[:, :, None]) ** 2 / s2)                   # (k, P, ws)
        img = np.matmul(gy, gx)
        img += rng.standard_normal(img.shape) * noise[:, None, None]
        dst[:] = img


def window_pairs(ws: int, n: int, seed: int, chunk: int = 1024, workers: int = 1):
    """Independent ``ws x ws`` float32 window pairs with a uniform random shift.

    Per window: shift ~ U[-8, 8]^2 (symmetric placement: frame A holds each
    particle at ``p - d/2``, frame B at ``p + d/2``), particle sigma
    ~ U(0.5, 1.2), peak intensities ~ U(0.5, 1), particles at density 0.05
    per px^2 over the window plus a 9 px margin, additive noise with sigma
    ~ U(0, 0.05) of the unit peak, unquantised.  Particles are separable
    Gaussians, so each window is the product ``(Gy diag(I0)) Gx`` and a chunk
    renders as one batched matrix product.  Chunk ``c`` draws from
    ``PCG64([seed, c])``, so the output does not depend on ``workers``
    (threads; NumPy releases the GIL in ``exp`` and ``matmul``).
    Returns ``(win_a, win_b, shift_uv)``, ``shift_uv[:, 0] = u`` (x).
    """
    wa = np.empty((n, ws, ws), np.float32)
    wb = np.empty((n, ws, ws), np.float32)
    shifts = np.empty((n, 2), np.float64)
    jobs = [(c, c * chunk, min(n, (c + 1) * chunk)) for c in range((n + chunk - 1) // chunk)]

    def run(job):
        c, c0, c1 = job
        _window_chunk(ws, seed, c, c1 - c0, wa[c0:c1], wb[c0:c1], shifts[c0:c1])

    if workers > 1:
        from concurrent.futures import ThreadPoolExecutor
        from threadpoolctl import threadpool_limits
        with threadpool_limits(1), ThreadPoolExecutor(workers) as ex:   # no nested BLAS threads
            list(ex.map(run, jobs))
    else:
        for job in jobs:
            run(job)
    return wa, wb, shifts


# Device renderer ------------------------------------------------------------

def particle_budget(shape, disp, model: ParticleModel):
    """``(margin, n_particles)`` of :func:`make_pair` for this field and model."""
    ht, wd = shape
    margin = int(math.ceil(getattr(disp, "max_abs", 0.0))) + int(math.ceil(4.0 * model.sigma)) + 1
    return margin, int(round(model.density * (ht + 2 * margin) * (wd + 2 * margin)))


def synth_field(disp):
    """C descriptor (``piv_synth_field``) of a displacement made by this module."""
    from . import _native
    p = disp.params
    f = _native.SynthField()
    if p["kind"] == "uniform":
        f.kind, f.u0, f.v0 = _native.FIELD_UNIFORM, p["u0"], p["v0"]
    elif p["kind"] == "sine":
        f.kind, f.amp_u, f.amp_v, f.period = _native.FIELD_SINE, p["amp_u"], p["amp_v"], p["period"]
    else:
        k = len(p["cx"])
        if k > _native.SYNTH_MAX_VORTICES:
            raise ValueError("too many vortices")
        f.kind, f.n_vortices, f.gamma = _native.FIELD_VORTICES, k, p["gamma"]
        for i in range(k):
            f.cx[i], f.cy[i], f.rc[i], f.circ[i] = p["cx"][i], p["cy"][i], p["rc"][i], p["circ"][i]
    return f


class DevicePairGenerator:
    """Renders seeded pairs of one frame shape directly into device buffers.

    Same particle model, margin, particle count, placement, noise, rounding
    and clipping as :func:`make_pair`; particles and noise come from a
    counter-based hash of ``seed`` instead of the host PCG64 stream, so a
    seed gives a different (equally distributed) realisation than
    :func:`make_pair`, identical on every run.
    """

    def __init__(self, shape, model: ParticleModel = None):
        from . import _native
        self._native = _native
        self.shape = tuple(int(t) for t in shape)
        self.model = model or ParticleModel()
        ht, wd = self.shape
        self.accum = _native.DeviceBuffer.zeros((2, ht, wd), np.int32)
        self.dtype = np.dtype(self.model.dtype)

    def new_frames(self):
        d = self._native.DeviceBuffer
        return d(self.shape, self.dtype), d(self.shape, self.dtype)

    def render(self, disp, seed: int, frame_a=None, frame_b=None, particles=None, stream=None):
        """Renders one pair; returns the two device buffers (allocated when
        not given).  ``particles``: optional device buffer (n, 5) FP32 that
        receives x, y, I0, du, dv of every particle."""
        import ctypes
        nat = self._native
        if frame_a is None or frame_b is None:
            frame_a, frame_b = self.new_frames()
        margin, n = particle_budget(self.shape, disp, self.model)
        m = nat.SynthModel()
        m.sigma, m.i0_low, m.i0_high = self.model.sigma, self.model.i0_low, self.model.i0_high
        m.noise_sigma, m.max_value = self.model.noise_sigma, self.model.max_value
        m.out_dtype, m.margin, m.n_particles = nat.in_dtype_code(self.dtype), margin, n
        f = synth_field(disp)
        ht, wd = self.shape
        nat.check(nat.load_library().piv_synth_pair(
            ctypes.byref(f), ctypes.byref(m), ctypes.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), ht, wd, frame_a.value,
            frame_b.value, self.accum.value, None if particles is None else particles.value,
            None if stream is None else stream.handle))
        return frame_a, frame_b

    @staticmethod
    def n_particles(shape, disp, model: ParticleModel):
        return particle_budget(shape, disp, model)[1]


def vortex_field(seed: int, shape):
    """The uneven-batch displacement field of seed ``seed`` (1-4 Lamb-Oseen
    vortices + shear), drawn exactly as :func:`cfg_vortex_uint16` draws it."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return lamb_oseen_displacement(rng, shape)


MODEL_12BIT = ParticleModel(sigma=0.65, density=0.05, i0_low=1000.0, i0_high=4095.0, noise_sigma=20.0,
                            max_value=4095.0, dtype=np.uint16)
