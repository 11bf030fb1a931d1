"""Throughput of the PIV hot path on one MI355X.

Workload (the metric's configuration): one 2048x2048 uint16 image pair with
12-bit particle images, passes 64 -> 32 -> 16 px at 50 % overlap, bilinear
window deformation between passes, 3x3 median validation (tol 2) with spring
replacement, GCV smoothn after the non-final passes, FP32.  A step is one
full pass loop on one pair whose frames are already resident in HBM; outputs
stay on the device.  Pairs are independent, so several are kept in flight:
one plan (stream, buffers, graph) per in-flight pair, steps assigned round
robin (--plans, default 3); the latency-bound field stages of one pair then
overlap the correlation kernels of another.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU
(pairs are independent; multi-GPU use shards the pair stream, no exchange).
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "openpiv-python-gpu_amd")
for _p in (PKG, ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "pairs/s at 2048², 64→32→16 px; learned vs plain at ≤0.05 px RMS error"

# Algorithmic FP32 FLOPs per window of the zero-padded FFT correlation
# (5 N log2 N per complex transform, forward row stage pruned for the zero
# padding): 12.5 F^2 log2 F + 6 F (F/2 + 1).
def fft_flops_per_window(F):
    return 12.5 * F * F * np.log2(F) + 6 * F * (F // 2 + 1)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stage_kernel_times(fa_u16, piv, reps=20):
    """Average duration of each pass's correlation launch and of the
    deformation launch, timed with HIP events on the stream the kernels are
    launched on."""
    from openpiv_gpu import _native
    lib = _native.load_library()
    stream = _native.Stream()
    sh = stream.handle
    ht, wd = fa_u16.shape
    d_u16 = _native.DeviceBuffer.from_array(fa_u16)
    d_f32 = _native.DeviceBuffer.from_array(fa_u16.astype(np.float32))
    desc = piv._plan_desc(_native.PIV_U16)

    def timed(launch):
        for _ in range(3):
            launch()
        e0, e1 = _native.Event(), _native.Event()
        e0.record(stream)
        for _ in range(reps):
            launch()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1) / reps

    out = []
    for k in range(desc.n_passes):
        p = desc.passes[k]
        m, n = p.field_m, p.field_n
        u = _native.DeviceBuffer((m, n), np.float32)
        v = _native.DeviceBuffer((m, n), np.float32)
        src, code = (d_u16, _native.PIV_U16) if k == 0 else (d_f32, _native.PIV_F32)

        def launch():
            _native.check(lib.piv_stage_correlate(src.ptr, src.ptr, code, ht, wd, ctypes.byref(p), None, None,
                                                  u.ptr, v.ptr, None, None, None, None, sh))
        ms = timed(launch)
        flops = fft_flops_per_window(p.fft_size) * m * n
        # search window <= F/2 without S2N: one lane per line for F <= 32, the
        # residue-split kernel for F = 64, 128 (launch_correlate's choice)
        half = 2 * p.search_size <= p.fft_size
        kname = ("corr_line_kernel" if half and p.fft_size <= 32 else
                 "corr_split_kernel" if half and p.fft_size in (64, 128) else "corr_kernel")
        out.append({"pass": k, "kernel": f"{kname}<F={p.fft_size},{'u16' if k == 0 else 'f32'}>",
                    "windows": m * n, "fft": p.fft_size,
                    "avg_ms": ms, "flop": flops, "tflops": flops / (ms * 1e-3) / 1e12})
    # deformation (second pass grid): reads both uint16 frames, writes two FP32 frames
    p = desc.passes[min(1, desc.n_passes - 1)]
    up = _native.DeviceBuffer.from_array(np.full((p.field_m, p.field_n), 1.5, np.float32))
    oa = _native.DeviceBuffer((ht, wd), np.float32)
    ob = _native.DeviceBuffer((ht, wd), np.float32)

    def dlaunch():
        _native.check(lib.piv_stage_deform(d_u16.ptr, d_u16.ptr, _native.PIV_U16, ht, wd, ctypes.byref(p),
                                           up.ptr, up.ptr, oa.ptr, ob.ptr, sh))
    ms = timed(dlaunch)
    nbytes = ht * wd * (2 * 2 + 2 * 4)
    deform = {"kernel": "deform_tile_kernel<u16>", "avg_ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9}
    return out, deform


def learned_vs_plain(n_pairs, weights_path, n_plans=3):
    """Plain 64->32->16 schedule vs the learned predictor (last pass only,
    full-schedule fallback above 5 % rejections) on a stream of held-out
    2048x2048 vortex pairs (seeds 200000+), rendered into HBM before timing.
    The headline comparison queues the pairs round robin over n_plans plans
    of each kind (piv_run_async; a learned plan runs its fallbacks when it is
    synchronised), as the plain headline does; the "synchronous" figures
    include each call's end-of-pair synchronisation."""
    from openpiv_gpu import _native, synthetic
    from openpiv_gpu.gpu_process import PIVGPU
    from openpiv_gpu.learned import CoarseFlowNet
    if not os.path.exists(weights_path):
        return {"skipped": f"no trained weights at {os.path.relpath(weights_path, ROOT)}"}
    shape = (2048, 2048)
    net = CoarseFlowNet.load(weights_path)
    kw = dict(search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float32")
    plain = PIVGPU(shape, 16, **kw)
    learned = PIVGPU(shape, 16, predictor=net, max_reject_fraction=0.05, **kw)
    # the network's field feeding the 32 px pass instead (only the 64 px pass skipped)
    learned32 = PIVGPU(shape, 16, predictor=net, max_reject_fraction=0.05, predictor_pass=1, **kw)
    # the network's field feeding two 16 px passes (schedule 64->32->16->16, the
    # first two skipped; the fallback runs all four)
    kw2 = dict(kw, search_size_iters=(1, 1, 2))
    learned16x2 = PIVGPU(shape, 16, predictor=net, max_reject_fraction=0.05, predictor_pass=2, **kw2)
    pp = plain._get_plan(_native.PIV_U16)
    lp = learned._get_plan(_native.PIV_U16)
    lp32 = learned32._get_plan(_native.PIV_U16)
    lp16x2 = learned16x2._get_plan(_native.PIV_U16)
    m, n = plain.piv_fields[-1].field_shape
    x, y = (np.asarray(c, np.float64) for c in plain.coords)
    frames = _native.DeviceBuffer((n_pairs, 2) + shape, np.uint16)
    outs = {k: _native.DeviceBuffer((n_pairs, 2, m, n), np.float32)
            for k in ("plain", "learned", "learned32", "learned16x2", "plain_q", "learned_q")}
    # queued plans: one network instance per learned plan (activations are per network)
    q_objs = {"plain_q": [PIVGPU(shape, 16, **kw) for _ in range(n_plans)],
              "learned_q": [PIVGPU(shape, 16, predictor=CoarseFlowNet.load(weights_path), max_reject_fraction=0.05,
                                   **kw) for _ in range(n_plans)]}
    q_plans = {k: [o._get_plan(_native.PIV_U16) for o in objs] for k, objs in q_objs.items()}
    gen = synthetic.DevicePairGenerator(shape, synthetic.MODEL_12BIT)
    seeds = [200000 + i for i in range(n_pairs)]
    log(f"rendering {n_pairs} held-out vortex pairs")
    for i, sd in enumerate(seeds):
        f = frames.view(i)
        gen.render(synthetic.vortex_field(sd, shape), sd, f.view(0), f.view(1))
    _native.synchronize()

    def run(plan, key, i):
        f = frames.view(i)
        o = outs[key].view(i)
        plan.run(f.view(0).value, f.view(1).value, True, o.view(0).value, o.view(1).value, None, None, True)

    res = {}
    for key, plan in (("plain", pp), ("learned", lp), ("learned32", lp32), ("learned16x2", lp16x2)):
        for w in range(3):
            run(plan, key, w % n_pairs)
        fallbacks = 0
        t0 = time.perf_counter()
        for i in range(n_pairs):
            run(plan, key, i)
            if key != "plain":
                fallbacks += plan.last_fallback()[0]
        dt = time.perf_counter() - t0
        res[key] = {"pairs_per_s": n_pairs / dt, "fallbacks": fallbacks}

    def run_async(plan, key, i):
        f = frames.view(i)
        o = outs[key].view(i)
        plan.run_async(f.view(0).value, f.view(1).value, o.view(0).value, o.view(1).value)

    for key, plans in q_plans.items():
        for w in range(3 * len(plans)):  # eager run and graph capture of every plan
            run_async(plans[w % len(plans)], key, w % n_pairs)
        for p in plans:
            p.synchronize()
        before = sum(p.learned_counts()[1] for p in plans)
        t0 = time.perf_counter()
        for i in range(n_pairs):
            run_async(plans[i % len(plans)], key, i)
        for p in plans:
            p.synchronize()
        dt = time.perf_counter() - t0
        res[key] = {"pairs_per_s": n_pairs / dt, "fallbacks": sum(p.learned_counts()[1] for p in plans) - before}
    for key in res:
        se = 0.0
        cnt = 0
        per_pair = []
        for i, sd in enumerate(seeds):
            uv = outs[key].view(i).get()
            ut, vt = synthetic.vortex_field(sd, shape)(x, y)
            e = float(np.sum((uv[0] - ut) ** 2 + (uv[1] - vt) ** 2))
            se += e
            cnt += 2 * ut.size
            per_pair.append(np.sqrt(e / (2 * ut.size)))
        res[key]["rms_px_vs_truth"] = float(np.sqrt(se / cnt))
        res[key]["worst_pair_rms_px"] = float(np.max(per_pair))
    return {
        "workload": f"{n_pairs} held-out 2048x2048 uint16 pairs (1-4 Lamb-Oseen vortices + shear, seeds 200000+), "
                    "device-rendered, resident in HBM",
        "submission": f"queued (piv_run_async) round robin over {n_plans} plans of each kind",
        "plain_pairs_per_s": round(res["plain_q"]["pairs_per_s"], 2),
        "learned_pairs_per_s": round(res["learned_q"]["pairs_per_s"], 2),
        "speedup": round(res["learned_q"]["pairs_per_s"] / res["plain_q"]["pairs_per_s"], 3),
        "plain_rms_px_vs_truth": round(res["plain_q"]["rms_px_vs_truth"], 5),
        "learned_rms_px_vs_truth": round(res["learned_q"]["rms_px_vs_truth"], 5),
        "learned_worst_pair_rms_px": round(res["learned_q"]["worst_pair_rms_px"], 5),
        "fallback_fraction": round(res["learned_q"]["fallbacks"] / n_pairs, 4),
        "max_reject_fraction": 0.05,
        "synchronous": {
            "plain_pairs_per_s": round(res["plain"]["pairs_per_s"], 2),
            "learned_pairs_per_s": round(res["learned"]["pairs_per_s"], 2),
            "speedup": round(res["learned"]["pairs_per_s"] / res["plain"]["pairs_per_s"], 3),
            "learned_rms_px_vs_truth": round(res["learned"]["rms_px_vs_truth"], 5),
            "fallback_fraction": round(res["learned"]["fallbacks"] / n_pairs, 4),
        },
        "two_16px_passes": {
            "submission": "synchronous",
            "pairs_per_s": round(res["learned16x2"]["pairs_per_s"], 2),
            "speedup": round(res["learned16x2"]["pairs_per_s"] / res["plain"]["pairs_per_s"], 3),
            "rms_px_vs_truth": round(res["learned16x2"]["rms_px_vs_truth"], 5),
            "worst_pair_rms_px": round(res["learned16x2"]["worst_pair_rms_px"], 5),
            "fallback_fraction": round(res["learned16x2"]["fallbacks"] / n_pairs, 4),
        },
        "skip_64_only": {
            "submission": "synchronous",
            "pairs_per_s": round(res["learned32"]["pairs_per_s"], 2),
            "speedup": round(res["learned32"]["pairs_per_s"] / res["plain"]["pairs_per_s"], 3),
            "rms_px_vs_truth": round(res["learned32"]["rms_px_vs_truth"], 5),
            "fallback_fraction": round(res["learned32"]["fallbacks"] / n_pairs, 4),
        },
        "weights": os.path.relpath(weights_path, ROOT),
    }


def uneven_batch(n_pairs, reps):
    """The uneven-batch workload: n_pairs device-rendered uint16 pairs (seed
    4000+i) of sizes {512^2, 1024x768, 2048^2, 3072x2048, 4096^2}, 1-4
    Lamb-Oseen vortices + shear, schedule 64->32->16 or 48->24 px, overlap
    U(0.5, 0.75); batched (one deformation / correlation launch per pass level,
    graph-captured, autotuned) vs one plan per pair."""
    from openpiv_gpu import _native, synthetic
    from openpiv_gpu.batch import PIVBatch, uneven_pair_spec
    from openpiv_gpu.gpu_process import PIVGPU
    specs = [uneven_pair_spec(4000 + i) for i in range(n_pairs)]
    gens = {}
    frames = []
    truth = []
    log(f"rendering {n_pairs} uneven pairs")
    for i, (shape, mss, kw) in enumerate(specs):
        if shape not in gens:
            gens[shape] = synthetic.DevicePairGenerator(shape, synthetic.MODEL_12BIT)
        disp = synthetic.vortex_field(4000 + i, shape)
        frames.append(gens[shape].render(disp, 4000 + i))
        truth.append(disp)
    _native.synchronize()
    log("building the batch")
    t0 = time.perf_counter()
    batch = PIVBatch(specs)
    batch.set_frames(frames)
    t_build = time.perf_counter() - t0
    batch.run()                    # eager
    choices = batch.tune(repeats=2, cache=False)
    batch.run()                    # capture + first replay
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.run()
    dt_batch = (time.perf_counter() - t0) / reps
    log("one plan per pair")
    pivs = [PIVGPU(shape, mss, dtype_f="float32", **kw) for shape, mss, kw in specs]
    plans = [p._get_plan(_native.PIV_U16) for p in pivs]
    outs = []
    for p, piv in zip(plans, pivs):
        m, n = piv.piv_fields[-1].field_shape
        outs.append((_native.DeviceBuffer((m, n), np.float32), _native.DeviceBuffer((m, n), np.float32)))

    def seq():
        for pl, (fa, fb), (u, v) in zip(plans, frames, outs):
            pl.run(fa.value, fb.value, True, u.value, v.value, None, None, True)
    seq()
    seq()
    t0 = time.perf_counter()
    for _ in range(reps):
        seq()
    dt_seq = (time.perf_counter() - t0) / reps
    se = cnt = 0.0
    for i, piv in enumerate(pivs):
        u, v, _, _ = batch.result(i)
        x, y = (np.asarray(c, np.float64) for c in piv.coords)
        ut, vt = truth[i](x, y)
        se += float(np.sum((u - ut) ** 2 + (v - vt) ** 2))
        cnt += 2 * u.size
    px = sum(s[0][0] * s[0][1] for s in specs)
    return {
        "metric": "pairs/s (uneven batch)", "value": round(n_pairs / dt_batch, 2), "unit": "pairs/s",
        "higher_is_better": True, "n_gpus": 1, "steps": reps, "data": "synthetic", "dtype": "f32",
        "config": {"workload": f"{n_pairs} uneven uint16 pairs (512^2..4096^2, 64->32->16 or 48->24 px, "
                               "overlap 0.5-0.75, Lamb-Oseen vortices + shear), frames resident in HBM",
                   "megapixels_per_batch": round(px / 1e6, 1)},
        "ms_per_batch": round(dt_batch * 1e3, 2),
        "one_plan_per_pair_pairs_per_s": round(n_pairs / dt_seq, 2),
        "batched_vs_per_pair": round(dt_seq / dt_batch, 3),
        "megapixel_pairs_per_s": round(px / 1e6 / dt_batch, 1),
        "rms_px_vs_truth": round(float(np.sqrt(se / cnt)), 5),
        "autotuned_choices": [int(c) for c in choices[:12]],
        "batch_build_s": round(t_build, 2),
    }


def tiff_stream(fa_dev_pairs, n_pairs, workdir):
    """Host-to-host rate of a stream of 2048x2048 uint16 TIFF pairs on disk:
    library TIFF decoding into pinned buffers, upload on a copy stream
    overlapped with the previous pair's compute, results back to the host;
    beside it the sequential host path (decode, piv_gpu on NumPy frames)."""
    from openpiv_gpu.gpu_process import PIVGPU
    from openpiv_gpu.io import PairStream, read_tiff
    try:
        from PIL import Image
    except ImportError:
        return {"skipped": "PIL not importable (needed to write the TIFF files)"}
    os.makedirs(workdir, exist_ok=True)
    files = []
    for k, (a, b) in enumerate(fa_dev_pairs):
        pa, pb = os.path.join(workdir, f"{k}_a.tif"), os.path.join(workdir, f"{k}_b.tif")
        Image.frombytes("I;16", a.shape[::-1], a.tobytes()).save(pa)
        Image.frombytes("I;16", b.shape[::-1], b.tobytes()).save(pb)
        files.append((pa, pb))
    seq = [files[i % len(files)] for i in range(n_pairs)]
    kw = dict(search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float32")
    stream = PairStream(PIVGPU((2048, 2048), 16, **kw), depth=3, workers=4)
    for _ in stream.run(seq[:4]):
        pass
    t0 = time.perf_counter()
    for _ in stream.run(seq):
        pass
    t_stream = time.perf_counter() - t0
    piv = PIVGPU((2048, 2048), 16, **kw)
    piv.run_host(read_tiff(seq[0][0]), read_tiff(seq[0][1]))
    t0 = time.perf_counter()
    for pa, pb in seq:
        piv.run_host(read_tiff(pa), read_tiff(pb))
    t_seq = time.perf_counter() - t0
    return {"pairs": n_pairs, "stream_pairs_per_s": round(n_pairs / t_stream, 2),
            "sequential_host_pairs_per_s": round(n_pairs / t_seq, 2),
            "note": "2048x2048 uint16 TIFF pairs from disk (page cache), decode + PCIe + compute + results to host"}


def variant_rates(fa, fb, disp, steps):
    """Device-resident pairs/s of the same 2048x2048 pair with a frame mask
    (disk of radius 300 px, 6.7 % of the frame) and with FP64 processing,
    plus their RMS against the known displacement (outside the mask)."""
    from openpiv_gpu import _native
    from openpiv_gpu.gpu_process import PIVGPU
    yy, xx = np.mgrid[0:2048, 0:2048]
    mask = (yy - 1000) ** 2 + (xx - 900) ** 2 < 300 ** 2
    out = {}
    for name, kw in (("masked", dict(mask=mask)), ("fp64", dict(dtype_f="float64"))):
        piv = PIVGPU((2048, 2048), 16, search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2,
                     **({"dtype_f": "float32"} | kw))
        plan = piv._get_plan(_native.PIV_U16)
        m, n = piv.piv_fields[-1].field_shape
        fd = piv.dtype_f
        da, db = _native.DeviceBuffer.from_array(fa), _native.DeviceBuffer.from_array(fb)
        u, v = _native.DeviceBuffer((m, n), fd), _native.DeviceBuffer((m, n), fd)
        for _ in range(3):
            plan.run(da.value, db.value, True, u.value, v.value, None, None, True)
        _native.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            plan.run(da.value, db.value, True, u.value, v.value, None, None, True)
        _native.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        x, y = piv.coords
        ut, vt = disp(np.asarray(x, np.float64), np.asarray(y, np.float64))
        uh, vh = u.get().astype(np.float64), v.get().astype(np.float64)
        ok = np.isfinite(uh) & np.isfinite(vh)
        rms = float(np.sqrt(np.mean((uh[ok] - ut[ok]) ** 2 + (vh[ok] - vt[ok]) ** 2) / 2))
        rec = {"pairs_per_s": round(1e3 / ms, 2), "ms_per_pair": round(ms, 4), "rms_px_vs_truth": round(rms, 5)}
        if name == "masked":
            rec["masked_vectors"] = int((~ok).sum())
            rec["smoothing_iterations"] = [plan.last_smoothing_iterations(k) for k in range(2)]
        out[name] = rec
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(fa, fb, timed=5):
    """The reference pipeline on the host's cores: the NumPy/SciPy oracle
    (FP32 correlation, FP64 smoothn) with scipy.fft and BLAS using every
    core this process is given; one warm-up pair, then the median of `timed`
    pairs, for the 2048x2048 64->32->16 workload and the 512x512 uint8
    single-pass one."""
    import scipy.fft
    from threadpoolctl import threadpool_limits
    from openpiv_gpu import synthetic
    from oracle import piv_oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    a1, b1, _ = synthetic.cfg_512_uint8(1)
    s2 = O.OracleSettings(min_search_size=16, search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2.0)
    s1 = O.OracleSettings(min_search_size=32, overlap_ratio=0.5, median_tol=2.0)

    def median_rate(a, b, s):
        times = []
        with scipy.fft.set_workers(cores), threadpool_limits(cores):
            O.piv(a, b, s)  # warm-up
            for _ in range(timed):
                t0 = time.perf_counter()
                O.piv(a, b, s)
                times.append(time.perf_counter() - t0)
        return 1.0 / float(np.median(times)), times

    r2, t2 = median_rate(fa, fb, s2)
    r1, t1 = median_rate(a1, b1, s1)
    return {"value": round(r2, 4), "unit": "pairs/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"2048x2048 uint16 64->32->16 workload: 1 warm-up pair, median of {timed} pairs "
                      f"({sum(t2):.1f} s timed); NumPy/SciPy oracle (FP32 correlation, FP64 smoothn), scipy.fft "
                      f"workers and BLAS threads = {cores}",
            "single_pass_512_uint8_pairs_per_s": round(r1, 3),
            "pair_seconds": [round(t, 3) for t in t2]}


def traffic_workload(steps=3, copies=3, copy_bytes=256 << 20):
    """The program the HBM-traffic counter runs profile: the 2048x2048 uint16
    64->32->16 pass loop (eager launches) on a device-rendered pair, then
    device-to-device copies of a known size that calibrate FETCH_SIZE and
    WRITE_SIZE."""
    from openpiv_gpu import _native, synthetic
    from openpiv_gpu.gpu_process import PIVGPU
    shape = (2048, 2048)
    gen = synthetic.DevicePairGenerator(shape, synthetic.MODEL_12BIT)
    fa, fb = gen.render(synthetic.sine_displacement(4.0, 3.0, 256.0), seed=2)
    piv = PIVGPU(shape, 16, search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float32",
                 use_graph=False)
    plan = piv._get_plan(_native.PIV_U16)
    m, n = piv.piv_fields[-1].field_shape
    u, v = _native.DeviceBuffer((m, n), np.float32), _native.DeviceBuffer((m, n), np.float32)
    for _ in range(steps):
        plan.run(fa.value, fb.value, True, u.value, v.value, None, None, True)
    src, dst = _native.DeviceBuffer((copy_bytes,), np.uint8), _native.DeviceBuffer((copy_bytes,), np.uint8)
    lib = _native.load_library()
    for _ in range(copies):
        _native.check(lib.piv_memcpy(dst.ptr, src.ptr, copy_bytes, 2))
    _native.synchronize()
    return {"workload": "traffic", "pass_loops": steps, "copies": copies, "copy_bytes": copy_bytes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--skip-cpu-baseline", action="store_true")
    ap.add_argument("--skip-kernel-timing", action="store_true")
    ap.add_argument("--learned-pairs", type=int, default=1000)
    ap.add_argument("--skip-learned", action="store_true")
    ap.add_argument("--workload", choices=("pair2048", "uneven", "traffic"), default="pair2048")
    ap.add_argument("--stream-pairs", type=int, default=48)
    ap.add_argument("--pairs", type=int, default=256)
    ap.add_argument("--variant-steps", type=int, default=20)
    ap.add_argument("--pairs-per-step", type=int, default=12,
                    help="pairs in one step of the headline loop (a batch of pairs spread over the plans)")
    ap.add_argument("--plans", type=int, default=3,
                    help="pairs in flight: one plan (stream) each, pairs assigned round robin")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for {max(args.gpus, world)}): image pairs are independent, "
              "so several GPUs would each run this same loop on their share of the pair stream.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "uneven":
        print(json.dumps(uneven_batch(args.pairs, max(1, args.steps))), flush=True)
        return
    if args.workload == "traffic":
        print(json.dumps(traffic_workload()), flush=True)
        return

    from openpiv_gpu import _native, synthetic
    from openpiv_gpu.gpu_process import PIVGPU

    log("generating the 2048x2048 uint16 pair (seed 2)")
    fa, fb, disp = synthetic.cfg_2048_uint16(seed=2)
    n_plans = max(1, args.plans)
    pivs, plans, slots = [], [], []
    for _ in range(n_plans):
        # one plan per in-flight pair: its own stream, buffers and graph
        p_ = PIVGPU((2048, 2048), 16, search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2,
                    dtype_f="float32", use_graph=not args.no_graph)
        pl = p_._get_plan(_native.PIV_U16)
        m, n = p_.piv_fields[-1].field_shape
        bufs = (_native.DeviceBuffer.from_array(fa), _native.DeviceBuffer.from_array(fb),
                _native.DeviceBuffer((m, n), np.float32), _native.DeviceBuffer((m, n), np.float32),
                _native.DeviceBuffer((m, n), np.uint8), _native.DeviceBuffer((m, n), np.uint8))
        pivs.append(p_)
        plans.append(pl)
        slots.append((bufs, [b.value for b in bufs]))
    piv, plan = pivs[0], plans[0]
    ptrs = slots[0][1]
    u, v = slots[0][0][2], slots[0][0][3]

    def sync_all():
        for pl in plans:
            pl.synchronize()
        _native.synchronize()

    pps = max(1, args.pairs_per_step)
    log(f"warmup {args.warmup} steps of {pps} pairs")
    for pl, (_, p) in zip(plans, slots):
        pl.run(p[0], p[1], True, p[2], p[3], p[4], p[5], True)  # first call: eager run, then the graph
    for k in range(args.warmup * pps):
        p = slots[k % n_plans][1]
        plans[k % n_plans].run_async(p[0], p[1], p[2], p[3], p[4], p[5])
    sync_all()
    log(f"timing {args.steps} steps of {pps} pairs over {n_plans} concurrent plans")
    t0 = time.perf_counter()
    for k in range(args.steps * pps):
        # each pair queued on the next plan's stream (round robin); the
        # timed region ends when every plan's stream has drained
        p = slots[k % n_plans][1]
        plans[k % n_plans].run_async(p[0], p[1], p[2], p[3], p[4], p[5])
    sync_all()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = 1e3 * pps / ms_per_step
    # one plan alone: pairs queued back to back on its stream
    t0 = time.perf_counter()
    for _ in range(args.steps * pps):
        plan.run_async(ptrs[0], ptrs[1], ptrs[2], ptrs[3], ptrs[4], ptrs[5])
    sync_all()
    single_plan_pairs_per_s = args.steps * pps / (time.perf_counter() - t0)
    # the synchronous call (returns after each pair's outputs are written)
    t0 = time.perf_counter()
    for _ in range(args.steps * pps):
        plan.run(ptrs[0], ptrs[1], True, ptrs[2], ptrs[3], ptrs[4], ptrs[5], True)
    sync_pairs_per_s = args.steps * pps / (time.perf_counter() - t0)

    # accuracy of this run against the known displacement
    x, y = piv.coords
    ut, vt = disp(np.asarray(x, np.float64), np.asarray(y, np.float64))
    uh, vh = u.get(), v.get()
    rms = float(np.sqrt(np.mean((uh - ut) ** 2 + (vh - vt) ** 2) / 2))

    # host-to-host rate (NumPy in/out, includes PCIe both ways)
    t2 = time.perf_counter()
    h2h_n = 5
    for _ in range(h2h_n):
        piv.run_host(fa, fb)
    h2h = h2h_n / (time.perf_counter() - t2)

    roof = None
    kernels = None
    deform = None
    if not args.skip_kernel_timing:
        log("per-kernel timing")
        kernels, deform = stage_kernel_times(fa, piv)
        dom = max(kernels, key=lambda r: r["avg_ms"])
        traffic = None
        tp = os.path.join(ROOT, "profiles", "hbm_traffic_2048_pipeline.json")
        if os.path.exists(tp):
            with open(tp) as f:
                tk = json.load(f)["kernels"].get(dom["kernel"])
            traffic = None if tk is None else tk["traffic_bytes"]
        roof = {"bound": "valu-fp32", "compute": "FP32 vector ALU (no matrix cores); FP32 vector peak",
                "kernel": dom["kernel"], "achieved": round(dom["tflops"], 3), "peak": 157.3, "unit": "TFLOP/s",
                "frac": round(dom["tflops"] / 157.3, 4), "traffic": traffic,
                "traffic_source": "profiles/hbm_traffic_2048_pipeline.md (FETCH_SIZE x2 + WRITE_SIZE, copy-calibrated)",
                "per_launch_flop": dom["flop"], "avg_launch_ms": round(dom["avg_ms"], 5)}

    stream = None
    if args.stream_pairs > 0:
        log("TIFF pair stream")
        import tempfile
        with tempfile.TemporaryDirectory() as tmp:
            stream = tiff_stream([(fa, fb)], args.stream_pairs, tmp)

    variants = None
    if args.variant_steps > 0:
        log("masked and FP64 variants")
        variants = variant_rates(fa, fb, disp, args.variant_steps)

    learned = None
    if not args.skip_learned and args.learned_pairs > 0:
        from openpiv_gpu.learned import default_weights_path
        log("learned predictor vs plain")
        learned = learned_vs_plain(args.learned_pairs, default_weights_path((2048, 2048)), n_plans=args.plans)

    cpu = None
    if not args.skip_cpu_baseline:
        log("CPU baseline (oracle on the host cores)")
        cpu = cpu_baseline(fa, fb)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2048x2048 uint16 pair (12-bit particles), passes 64->32->16 px, 50% overlap, "
                               "deformation + median(2) + spring + GCV smoothn, plain method",
                   "pairs_per_step": pps, "graph": not args.no_graph,
                   "submission": f"queued (piv_run_async) round robin over {n_plans} plans, one stream each",
                   "concurrent_plans": n_plans},
        "rms_px_vs_truth": round(rms, 5),
        "single_plan_pairs_per_s": round(single_plan_pairs_per_s, 2),
        "sync_call_pairs_per_s": round(sync_pairs_per_s, 2),
        "host_to_host_pairs_per_s": round(h2h, 2),
        "roofline": roof,
        "kernels": kernels,
        "deform_hbm": None if deform is None else {
            "bound": "hbm", "achieved": round(deform["gbs"], 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(deform["gbs"] / 8000.0, 4), "avg_launch_ms": round(deform["avg_ms"], 5),
            "per_launch_bytes": deform["bytes"]},
        "learned_vs_plain": learned,
        "variants": variants,
        "tiff_stream": stream,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
