"""Precision sweep of the window correlation: the FP32 FFT path and the DFT
matrix-product path on the MFMA units (FP16/BF16 operands, plain or split
hi/lo, FP32 accumulation) against an FP64 FFT correlation of the same
windows.

Workload: ``--n`` independent window pairs per size (default 65,536 at 16,
32 and 64 px), each with a uniform random shift in [-8, 8]^2 px, particles at
0.05 / px^2 with sigma ~ U(0.5, 1.2) px and noise sigma ~ U(0, 5 %) of the
peak (``openpiv_gpu.synthetic.window_pairs``), correlated at F = 2 W with the
3-point Gaussian fit.

Reported per size and path:
  * ``rms_vs_fp64`` -- RMS of (u, v) minus the FP64 result over windows where
    both find the same integer peak, the FP64 plane has no near tie and no
    Gaussian-fit sample lies within 1e-4 (of the peak) of zero;
  * ``fit_decision_flips`` -- windows left out by that last rule whose fit
    decision (all five samples > 0) came out differently;
  * ``peak_agree`` -- fraction of windows with the FP64 integer peak;
  * ``rms_vs_truth`` -- RMS against the generated shift over windows the FP64
    path resolves within 1 px;
  * ``bias`` -- peak-locking curve: mean error against the generated shift,
    binned by the fractional part of the true shift (10 bins, u and v
    pooled), its peak-to-peak amplitude, and the largest bin difference to
    the FP64 curve;
  * wall time per path (host call, device-resident windows).

Run on a GPU host from the repository root:
    python tests/precision_sweep.py --out precision.json
"""

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (os.path.join(ROOT, "openpiv-python-gpu_amd"), ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402
import scipy.fft  # noqa: E402

from oracle import piv_oracle as O  # noqa: E402

MODES = ((0, "fp16_split"), (1, "bf16_split"), (2, "fp16"), (3, "bf16"))


def log(msg):
    print(msg, flush=True)


def fp64_reference(wa, wb, fsize, workers, chunk=4096):
    n = wa.shape[0]
    u = np.empty(n)
    v = np.empty(n)
    ip = np.empty(n, np.int64)
    jp = np.empty(n, np.int64)
    tie = np.empty(n, bool)
    amb = np.empty(n, bool)
    with scipy.fft.set_workers(workers):
        for c0 in range(0, n, chunk):
            sl = slice(c0, min(n, c0 + chunk))
            planes = O.correlate_windows(wa[sl].astype(np.float64), wb[sl].astype(np.float64), fsize)
            i, j = O.first_peak(planes)
            isp, jsp = O.subpixel_peak(planes, i, j)
            u[sl], v[sl], ip[sl], jp[sl] = jsp - fsize // 2, isp - fsize // 2, i, j
            flat = planes.reshape(planes.shape[0], -1)
            part = np.partition(flat, -2, axis=1)
            tie[sl] = (part[:, -1] - part[:, -2]) <= 1e-5 * np.maximum(np.abs(part[:, -1]), 1e-30)
            # a Gaussian-fit sample within 1e-4 of zero (relative to the peak):
            # whether the fit runs there depends on rounding
            ii = np.clip(i, 1, fsize - 2)
            jj = np.clip(j, 1, fsize - 2)
            k = np.arange(planes.shape[0])
            samples = np.stack([planes[k, ii, jj], planes[k, ii - 1, jj], planes[k, ii + 1, jj],
                                planes[k, ii, jj - 1], planes[k, ii, jj + 1]])
            amb[sl] = np.any(np.abs(samples) <= 1e-4 * np.abs(planes[k, ii, jj]), axis=0)
    return u, v, ip, jp, tie, amb


def bias_curve(err_u, err_v, true_u, true_v, ok, bins=10):
    frac = np.concatenate([(true_u - np.floor(true_u))[ok], (true_v - np.floor(true_v))[ok]])
    err = np.concatenate([err_u[ok], err_v[ok]])
    idx = np.minimum((frac * bins).astype(int), bins - 1)
    return np.array([err[idx == b].mean() for b in range(bins)])


def summarise(u, v, ip, jp, ref, shifts, resolved, secs):
    ru, rv, rip, rjp, tie, amb = ref
    same = (ip == rip) & (jp == rjp)
    good = same & ~tie & ~amb
    out = {
        "fit_decision_flips": int(np.sum(same & ~tie & amb & (np.maximum(np.abs(u - ru), np.abs(v - rv)) > 0.01))),
        "rms_vs_fp64": float(np.sqrt(np.mean(((u - ru) ** 2 + (v - rv) ** 2)[good]) / 2)),
        "max_abs_vs_fp64": float(np.max(np.maximum(np.abs(u - ru), np.abs(v - rv))[good])),
        "peak_agree": float(same.mean()),
        "rms_vs_truth": float(np.sqrt(np.mean(((u - shifts[:, 0]) ** 2 + (v - shifts[:, 1]) ** 2)[resolved]) / 2)),
        "seconds": round(secs, 4),
    }
    curve = bias_curve(u - shifts[:, 0], v - shifts[:, 1], shifts[:, 0], shifts[:, 1], resolved)
    out["bias_curve"] = [round(float(b), 6) for b in curve]
    out["bias_pp"] = float(curve.max() - curve.min())
    return out, curve


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--sizes", type=int, nargs="+", default=[16, 32, 64])
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()

    from openpiv_gpu import _native, synthetic
    import gpu_util as gu

    lib = _native.load_library()
    report = {"workload": f"{args.n} independent window pairs per size, shift U[-8,8]^2 px, density 0.05/px^2, "
                          f"sigma U(0.5,1.2) px, noise U(0,5%) of peak, F = 2W, Gaussian 3-point fit",
              "sizes": {}}
    for ws in args.sizes:
        fsize = 2 * ws
        t0 = time.perf_counter()
        wa, wb, shifts = synthetic.window_pairs(ws, args.n, seed=args.seed + ws, workers=args.workers)
        log(f"W={ws}: generated {args.n} pairs in {time.perf_counter() - t0:.1f} s")
        t0 = time.perf_counter()
        ref = fp64_reference(wa, wb, fsize, args.workers)
        log(f"W={ws}: FP64 reference in {time.perf_counter() - t0:.1f} s")
        ru, rv = ref[0], ref[1]
        resolved = (np.abs(ru - shifts[:, 0]) < 1) & (np.abs(rv - shifts[:, 1]) < 1)
        fp64_curve = bias_curve(ru - shifts[:, 0], rv - shifts[:, 1], shifts[:, 0], shifts[:, 1], resolved)
        entry = {"fft": fsize, "resolved_fraction": float(resolved.mean()),
                 "fp64": {"rms_vs_truth": float(np.sqrt(np.mean(((ru - shifts[:, 0]) ** 2 +
                                                                  (rv - shifts[:, 1]) ** 2)[resolved]) / 2)),
                          "bias_curve": [round(float(b), 6) for b in fp64_curve],
                          "bias_pp": float(fp64_curve.max() - fp64_curve.min())}}

        n = args.n
        da = gu.dev(wa)
        db = gu.dev(wb)
        u = gu.empty((n,), np.float32)
        v = gu.empty((n,), np.float32)
        ipd = gu.empty((n,), np.int32)
        jpd = gu.empty((n,), np.int32)
        geom = gu.stack_geometry(n, ws)
        desc = gu.pass_desc(geom, fsize)
        import ctypes

        def fft_call():
            _native.check(lib.piv_stage_correlate(gu.vp(da), gu.vp(db), _native.PIV_F32, n * ws, ws,
                                                  ctypes.byref(desc), None, None, gu.vp(u), gu.vp(v), gu.vp(ipd),
                                                  gu.vp(jpd), None, None, None))
            _native.synchronize()

        def dft_call(prec):
            _native.check(lib.piv_stage_correlate_dft(gu.vp(da), gu.vp(db), n, ws, fsize, prec, 0, gu.vp(u),
                                                      gu.vp(v), gu.vp(ipd), gu.vp(jpd), None, None))
            _native.synchronize()

        for name, call in [("fp32_fft", fft_call)] + [(nm, (lambda p=p: dft_call(p))) for p, nm in MODES]:
            call()  # warm (library tables, code objects)
            t0 = time.perf_counter()
            call()
            secs = time.perf_counter() - t0
            s, curve = summarise(u.get().astype(np.float64), v.get().astype(np.float64), ipd.get(), jpd.get(),
                                 ref, shifts, resolved, secs)
            s["bias_vs_fp64_max"] = float(np.abs(curve - fp64_curve).max())
            entry[name] = s
            log(f"W={ws} {name}: rms_vs_fp64={s['rms_vs_fp64']:.3e} agree={s['peak_agree']:.5f} "
                f"rms_truth={s['rms_vs_truth']:.4f} bias_pp={s['bias_pp']:.4f} "
                f"dbias={s['bias_vs_fp64_max']:.2e} t={secs * 1e3:.1f} ms")
        report["sizes"][str(ws)] = entry
    text = json.dumps(report, indent=1)
    if args.out:
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        with open(args.out, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()
