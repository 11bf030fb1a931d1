"""The committed, GPU-trained coarse-flow network on held-out pairs: the
learned predictor (the network's field replaces the 64 and 32 px passes as
the deformation predictor, gpu_process.py:619-628; one 16 px pass remains,
with the full schedule as fallback above 5 % rejections) against the plain
64 -> 32 -> 16 schedule and against the same 16 px pass without any
predictor, on 16 held-out 2048 x 2048 uint16 vortex pairs (seeds 200000+,
never used in training).

What the network gains is stated as tests: no pair falls back; the learned
path's RMS error against the true displacement is within 0.003 px of the
plain schedule's; the single 16 px pass without the network is several
times worse (displacements up to 10 px exceed what a 16 px window recovers
on its own); and the learned pair takes less time than the plain one."""

import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_PAIRS = 16
SHAPE = (2048, 2048)


@pytest.fixture(scope="module")
def runs():
    from openpiv_gpu import _native, synthetic
    from openpiv_gpu.gpu_process import PIVGPU
    from openpiv_gpu.learned import CoarseFlowNet
    path = os.path.join(os.path.dirname(synthetic.__file__), "weights", "coarse_flow_2048x2048.npz")
    assert os.path.exists(path), "committed weights missing"
    net = CoarseFlowNet.load(path)
    kw = dict(overlap_ratio=0.5, median_tol=2, dtype_f="float32")
    plans = {
        "plain": PIVGPU(SHAPE, 16, search_size_iters=(1, 1, 1), **kw),
        "learned": PIVGPU(SHAPE, 16, search_size_iters=(1, 1, 1), predictor=net, max_reject_fraction=0.05, **kw),
        "single16": PIVGPU(SHAPE, 16, search_size_iters=(1,), **kw),
    }
    raw = {k: p._get_plan(_native.PIV_U16) for k, p in plans.items()}
    m, n = plans["plain"].piv_fields[-1].field_shape
    assert plans["single16"].piv_fields[-1].field_shape == (m, n)
    x, y = (np.asarray(c, np.float64) for c in plans["plain"].coords)
    gen = synthetic.DevicePairGenerator(SHAPE, synthetic.MODEL_12BIT)
    frames = _native.DeviceBuffer((N_PAIRS, 2) + SHAPE, np.uint16)
    seeds = [200000 + i for i in range(N_PAIRS)]
    for i, sd in enumerate(seeds):
        f = frames.view(i)
        gen.render(synthetic.vortex_field(sd, SHAPE), sd, f.view(0), f.view(1))
    _native.synchronize()
    out = _native.DeviceBuffer((2, m, n), np.float32)
    res = {}
    for key, plan in raw.items():
        f = frames.view(0)
        for _ in range(2):  # eager run, graph capture
            plan.run(f.view(0).value, f.view(1).value, True, out.view(0).value, out.view(1).value, None, None, True)
        errs, fallbacks, secs = [], 0, 0.0
        for i, sd in enumerate(seeds):
            f = frames.view(i)
            t0 = time.perf_counter()
            plan.run(f.view(0).value, f.view(1).value, True, out.view(0).value, out.view(1).value, None, None, True)
            secs += time.perf_counter() - t0
            if key == "learned":
                fallbacks += plan.last_fallback()[0]
            uv = out.get()
            ut, vt = synthetic.vortex_field(sd, SHAPE)(x, y)
            errs.append(float(np.mean((uv[0] - ut) ** 2 + (uv[1] - vt) ** 2) / 2))
        res[key] = {"rms": float(np.sqrt(np.mean(errs))), "worst": float(np.sqrt(np.max(errs))),
                    "fallbacks": fallbacks, "ms_per_pair": 1e3 * secs / N_PAIRS}
    print(res, flush=True)
    return res


def test_learned_path_never_falls_back(runs):
    assert runs["learned"]["fallbacks"] == 0


def test_learned_rms_matches_plain_schedule(runs):
    assert runs["learned"]["rms"] <= runs["plain"]["rms"] + 0.003, runs
    assert runs["learned"]["worst"] <= 0.15, runs


def test_network_gain_over_the_pass_without_it(runs):
    # the same single 16 px pass with no predictor cannot follow the vortices
    assert runs["single16"]["rms"] >= 3 * runs["learned"]["rms"], runs


def test_learned_pair_is_faster_than_plain(runs):
    assert runs["learned"]["ms_per_pair"] < runs["plain"]["ms_per_pair"] / 1.15, runs
