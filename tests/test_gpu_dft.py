"""Precision mode: window correlation as DFT matrix products on the MFMA units
(FP16/BF16 inputs with split hi/lo operands, FP32 accumulation), checked
against the FP64 oracle and the FP32 FFT path on seeded window pairs."""

import numpy as np
import pytest

from oracle import piv_oracle as O

pytestmark = pytest.mark.gpu
gu = pytest.importorskip("gpu_util")


def _fp64_reference(wa, wb, fsize):
    planes = O.correlate_windows(wa.astype(np.float64), wb.astype(np.float64), fsize)
    ip, jp = O.first_peak(planes)
    isp, jsp = O.subpixel_peak(planes, ip, jp)
    return jsp - fsize // 2, isp - fsize // 2, ip, jp, planes


@pytest.mark.parametrize("ws", [16, 32, 64])
def test_dft_modes_against_fp64(ws):
    from openpiv_gpu import synthetic
    fsize = 2 * ws
    wa, wb, shifts = synthetic.window_pairs(ws, 512, seed=3 + ws)
    u64, v64, ip64, jp64, planes = _fp64_reference(wa, wb, fsize)
    ok = ~gu.near_tie_mask(planes) & ~gu.fit_sign_ambiguous(planes, ip64, jp64, rel=1e-3)
    fft = gu.stack_correlate_fft(wa, wb, fsize)
    rms_fft = np.sqrt(np.mean((fft["u"] - u64)[ok] ** 2 + (fft["v"] - v64)[ok] ** 2))
    assert rms_fft < 1e-4
    res = {}
    for prec, name in ((0, "fp16x2"), (1, "bf16x2"), (2, "fp16"), (3, "bf16")):
        got = gu.stack_correlate_dft(wa, wb, fsize, prec)
        same_peak = (got["i_peak"] == ip64) & (got["j_peak"] == jp64)
        good = ok & same_peak
        res[name] = (np.sqrt(np.mean((got["u"] - u64)[good] ** 2 + (got["v"] - v64)[good] ** 2)), same_peak[ok].mean())
    print(ws, {k: (float(a), float(b)) for k, (a, b) in res.items()}, float(rms_fft))
    # split operands recover most of FP32's accuracy; single half-precision
    # operands are coarser but still locate the integer peak
    assert res["fp16x2"][0] < 2e-3 and res["fp16x2"][1] > 0.99
    assert res["bf16x2"][0] < 2e-2 and res["bf16x2"][1] > 0.98
    assert res["fp16"][1] > 0.95 and res["bf16"][1] > 0.9


def test_dft_planes_match_fft_path():
    from openpiv_gpu import synthetic
    wa, wb, _ = synthetic.window_pairs(32, 64, seed=99)
    fft_planes = O.correlate_windows(wa.astype(np.float64), wb.astype(np.float64), 64)
    got = gu.stack_correlate_dft(wa, wb, 64, 0, planes=True)["corr"]
    # the DFT path scales each window and spectrum by its maximum: compare
    # shapes after normalising by the plane maximum
    a = fft_planes / np.abs(fft_planes).reshape(64, -1).max(1)[:, None, None]
    b = got / np.abs(got).reshape(64, -1).max(1)[:, None, None]
    assert np.abs(a - b).max() < 2e-3
