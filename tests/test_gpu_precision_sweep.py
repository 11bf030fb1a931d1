"""Precision sweep of the window correlation (the DFT-as-matmul workload):
the FP32 FFT path and the DFT matrix-product path on the MFMA units (FP16 /
BF16 operands, plain or split hi/lo, FP32 accumulation) against the FP64
oracle on the same windows.

Workload per window size W in {16, 32, 64}: independent window pairs
(``synthetic.window_pairs``: shift U[-8, 8]^2 px, particles at 0.05 / px^2,
sigma U(0.5, 1.2) px, noise U(0, 5 %) of the peak), correlated at F = 2W with
the 3-point Gaussian fit.  16,384 pairs per size by default; the environment
variable PIV_SWEEP_N sets another count (65,536 for the full sweep) and
PIV_SWEEP_REPORT a path for the JSON report.

Per size and path: ``rms_vs_fp64`` over windows where both find the same
integer peak, the FP64 plane has no near tie and no Gaussian-fit sample
lies within 1e-4 (of the peak) of zero; ``peak_agree``; ``rms_vs_truth``
over windows the FP64 path resolves within 1 px; the peak-locking curve
(mean error against the generated shift binned by its fractional part) and
its largest difference to the FP64 curve.  Asserted bounds are written in
``BOUNDS``."""

import ctypes
import json
import os
import time

import numpy as np
import pytest
import scipy.fft

from oracle import piv_oracle as O

pytestmark = pytest.mark.gpu
gu = pytest.importorskip("gpu_util")

MODES = ((0, "fp16_split"), (1, "bf16_split"), (2, "fp16"), (3, "bf16"))
# path -> (max rms_vs_fp64 px, min peak agreement, max peak-locking change px)
BOUNDS = {"fp32_fft": (1e-5, 0.9999, 1e-6), "fp16_split": (1e-5, 0.9999, 1e-6),
          "bf16_split": (1e-4, 0.999, 1e-5), "fp16": (5e-4, 0.999, 5e-3), "bf16": (5e-3, 0.995, 5e-3)}


def _fp64_reference(wa, wb, fsize, workers=16, chunk=4096):
    n = wa.shape[0]
    u, v = np.empty(n), np.empty(n)
    ip, jp = np.empty(n, np.int64), np.empty(n, np.int64)
    tie, amb = np.empty(n, bool), np.empty(n, bool)
    with scipy.fft.set_workers(workers):
        for c0 in range(0, n, chunk):
            sl = slice(c0, min(n, c0 + chunk))
            planes = O.correlate_windows(wa[sl].astype(np.float64), wb[sl].astype(np.float64), fsize)
            i, j = O.first_peak(planes)
            isp, jsp = O.subpixel_peak(planes, i, j)
            u[sl], v[sl], ip[sl], jp[sl] = jsp - fsize // 2, isp - fsize // 2, i, j
            tie[sl] = gu.near_tie_mask(planes)
            ii, jj = np.clip(i, 1, fsize - 2), np.clip(j, 1, fsize - 2)
            k = np.arange(planes.shape[0])
            samples = np.stack([planes[k, ii, jj], planes[k, ii - 1, jj], planes[k, ii + 1, jj],
                                planes[k, ii, jj - 1], planes[k, ii, jj + 1]])
            amb[sl] = np.any(np.abs(samples) <= 1e-4 * np.abs(planes[k, ii, jj]), axis=0)
    return u, v, ip, jp, tie, amb


def _bias_curve(u, v, shifts, ok, bins=10):
    tu, tv = shifts[:, 0], shifts[:, 1]
    frac = np.concatenate([(tu - np.floor(tu))[ok], (tv - np.floor(tv))[ok]])
    err = np.concatenate([(u - tu)[ok], (v - tv)[ok]])
    idx = np.minimum((frac * bins).astype(int), bins - 1)
    return np.array([err[idx == b].mean() for b in range(bins)])


def _summary(u, v, ip, jp, ref, shifts, resolved, fp64_curve, secs):
    ru, rv, rip, rjp, tie, amb = ref
    same = (ip == rip) & (jp == rjp)
    good = same & ~tie & ~amb
    curve = _bias_curve(u, v, shifts, resolved)
    return {"rms_vs_fp64": float(np.sqrt(np.mean(((u - ru) ** 2 + (v - rv) ** 2)[good]) / 2)),
            "max_abs_vs_fp64": float(np.max(np.maximum(np.abs(u - ru), np.abs(v - rv))[good])),
            "peak_agree": float(same[~tie].mean()),
            "rms_vs_truth": float(np.sqrt(np.mean(((u - shifts[:, 0]) ** 2 + (v - shifts[:, 1]) ** 2)[resolved])
                                          / 2)),
            "bias_pp": float(curve.max() - curve.min()),
            "bias_vs_fp64_max": float(np.abs(curve - fp64_curve).max()),
            "bias_curve": [round(float(b), 6) for b in curve],
            "compared": int(good.sum()), "seconds": round(secs, 4)}


@pytest.fixture(scope="module")
def report():
    rep = {"workload": "independent window pairs per size, shift U[-8,8]^2 px, density 0.05/px^2, sigma "
                       "U(0.5,1.2) px, noise U(0,5%) of peak, F = 2W, Gaussian 3-point fit", "sizes": {}}
    yield rep
    path = os.environ.get("PIV_SWEEP_REPORT")
    if path:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            json.dump(rep, f, indent=1)


@pytest.mark.parametrize("ws", [16, 32, 64])
def test_precision_sweep(ws, report):
    from openpiv_gpu import _native, synthetic
    lib = _native.load_library()
    n = int(os.environ.get("PIV_SWEEP_N", "16384"))
    fsize = 2 * ws
    wa, wb, shifts = synthetic.window_pairs(ws, n, seed=2024 + ws, workers=16)
    ref = _fp64_reference(wa, wb, fsize)
    ru, rv = ref[0], ref[1]
    resolved = (np.abs(ru - shifts[:, 0]) < 1) & (np.abs(rv - shifts[:, 1]) < 1)
    fp64_curve = _bias_curve(ru, rv, shifts, resolved)
    entry = {"n": n, "fft": fsize, "resolved_fraction": float(resolved.mean()),
             "fp64": {"rms_vs_truth": float(np.sqrt(np.mean(((ru - shifts[:, 0]) ** 2 +
                                                              (rv - shifts[:, 1]) ** 2)[resolved]) / 2)),
                      "bias_curve": [round(float(b), 6) for b in fp64_curve],
                      "bias_pp": float(fp64_curve.max() - fp64_curve.min())}}
    da, db = gu.dev(wa), gu.dev(wb)
    u, v = gu.empty((n,), np.float32), gu.empty((n,), np.float32)
    ipd, jpd = gu.empty((n,), np.int32), gu.empty((n,), np.int32)
    desc = gu.pass_desc(gu.stack_geometry(n, ws), fsize)

    def fft_call():
        _native.check(lib.piv_stage_correlate(gu.vp(da), gu.vp(db), _native.PIV_F32, n * ws, ws, ctypes.byref(desc),
                                              None, None, gu.vp(u), gu.vp(v), gu.vp(ipd), gu.vp(jpd), None, None,
                                              None))
        _native.check(lib.piv_synchronize())

    def dft_call(prec):
        _native.check(lib.piv_stage_correlate_dft(gu.vp(da), gu.vp(db), n, ws, fsize, prec, 0, gu.vp(u), gu.vp(v),
                                                  gu.vp(ipd), gu.vp(jpd), None, None))
        _native.check(lib.piv_synchronize())

    failures = []
    for name, call in [("fp32_fft", fft_call)] + [(nm, (lambda p=p: dft_call(p))) for p, nm in MODES]:
        call()
        t0 = time.perf_counter()
        call()
        secs = time.perf_counter() - t0
        s = _summary(u.get().astype(np.float64), v.get().astype(np.float64), ipd.get(), jpd.get(), ref, shifts,
                     resolved, fp64_curve, secs)
        entry[name] = s
        rms_max, agree_min, bias_max = BOUNDS[name]
        if not (s["rms_vs_fp64"] <= rms_max and s["peak_agree"] >= agree_min and s["bias_vs_fp64_max"] <= bias_max):
            failures.append((name, s["rms_vs_fp64"], s["peak_agree"], s["bias_vs_fp64_max"]))
        print(f"W={ws} {name}: rms_vs_fp64={s['rms_vs_fp64']:.3e} agree={s['peak_agree']:.6f} "
              f"dbias={s['bias_vs_fp64_max']:.2e} rms_truth={s['rms_vs_truth']:.4f} t={secs * 1e3:.1f} ms",
              flush=True)
    report["sizes"][str(ws)] = entry
    assert not failures, failures
