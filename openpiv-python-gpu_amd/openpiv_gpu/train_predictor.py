"""Trains the learned coarse predictor on the GPU.

Training pairs: seeded 2048 x 2048 uint16 pairs of the uneven-batch
generator (1-4 Lamb-Oseen vortices + shear, 12-bit particle model), rendered
on the device (``synthetic.DevicePairGenerator``).  Target: the field that
the plain 64 -> 32 -> 16 px schedule (50 % overlap, median 2 + spring,
smoothn between passes) returns for the pair, on the network's output nodes
(every second node of the 16 px grid), weighted 0 where that pass's median
test rejected the vector.  The network input is the pair box-averaged 8 x 8
and standardised.  Every step runs the library's own forward, backward and
Adam kernels.

    python -m openpiv_gpu.train_predictor --pairs 20000 --epochs 8 --out weights.npz
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import _native, synthetic
from .gpu_process import PIVGPU
from .learned import FACTOR, CoarseFlowNet, default_weights_path

SHAPE = (2048, 2048)
TRAIN_SEED0 = 100000
HOLDOUT_SEED0 = 200000


def log(msg):
    print(msg, flush=True)


def plain_piv(shape=SHAPE, use_graph=True):
    return PIVGPU(shape, 16, search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float32",
                  use_graph=use_graph)


def node_truth(disp, out_shape):
    """The generating field on the network's output nodes (16 j + 8, 16 i + 8)."""
    ys = np.arange(out_shape[0]) * 16.0 + 8.0
    xs = np.arange(out_shape[1]) * 16.0 + 8.0
    xx, yy = np.meshgrid(xs, ys)
    u, v = disp(xx, yy)
    return np.stack([u, v]).astype(np.float32)


class Dataset:
    """Network inputs (device, BF16), plain-method targets and weights
    (device, FP32) and the generating field on the nodes (host) for a run of
    consecutive seeds."""

    def __init__(self, net: CoarseFlowNet, n_pairs: int, seed0: int, progress_every: int = 1000):
        self.n = n_pairs
        ih, iw = net.in_shape
        oh, ow = net.out_shape
        self.in_elems = ih * iw * 2
        self.out_elems = oh * ow
        self.inputs = _native.DeviceBuffer((n_pairs, ih, iw, 2), np.uint16)
        targets = np.empty((n_pairs, 2, oh, ow), np.float32)
        weights = np.empty((n_pairs, oh, ow), np.float32)
        self.truth = np.empty((n_pairs, 2, oh, ow), np.float32)
        gen = synthetic.DevicePairGenerator(net.frame_shape, synthetic.MODEL_12BIT)
        fa, fb = gen.new_frames()
        piv = plain_piv(net.frame_shape)
        plan = piv._get_plan(_native.PIV_U16)
        m, n = piv.piv_fields[-1].field_shape
        u = np.empty((m, n), np.float32)
        v = np.empty((m, n), np.float32)
        vi = np.empty((m, n), np.uint8)
        t0 = time.perf_counter()
        for i in range(n_pairs):
            seed = seed0 + i
            disp = synthetic.vortex_field(seed, net.frame_shape)
            gen.render(disp, seed, fa, fb)
            net.prepare_input(fa, fb, self.inputs.value + i * self.in_elems * 2)
            plan.run(fa.value, fb.value, True, u.ctypes.data, v.ctypes.data, vi.ctypes.data, None, False)
            targets[i, 0] = u[0::2, 0::2][:oh, :ow]
            targets[i, 1] = v[0::2, 0::2][:oh, :ow]
            weights[i] = (vi[0::2, 0::2][:oh, :ow] == 0)
            self.truth[i] = node_truth(disp, (oh, ow))
            if progress_every and (i + 1) % progress_every == 0:
                log(f"  dataset {i + 1}/{n_pairs} pairs ({time.perf_counter() - t0:.1f} s)")
        self.targets = _native.DeviceBuffer.from_array(targets)
        self.weights = _native.DeviceBuffer.from_array(weights)
        self.targets_host = targets
        self.weights_host = weights

    def input_ptr(self, i):
        return self.inputs.value + i * self.in_elems * 2

    def target_ptr(self, i):
        return self.targets.value + i * 2 * self.out_elems * 4

    def weight_ptr(self, i):
        return self.weights.value + i * self.out_elems * 4


def evaluate(net: CoarseFlowNet, data: Dataset, batch: int):
    """RMS (per component) of the network's field against the plain-method
    target (where not rejected) and against the generating field."""
    oh, ow = net.out_shape
    out = _native.DeviceBuffer((batch, 2, oh, ow), np.float32)
    se_t = se_g = wsum = 0.0
    for i0 in range(0, data.n - batch + 1, batch):
        net.forward(data.input_ptr(i0), batch, out.value)
        o = out.get()
        t = data.targets_host[i0:i0 + batch]
        w = data.weights_host[i0:i0 + batch][:, None]
        se_t += float(np.sum(w * (o - t) ** 2))
        wsum += float(np.sum(w)) * 2
        se_g += float(np.sum((o - data.truth[i0:i0 + batch]) ** 2))
    n_all = (data.n // batch) * batch * 2 * oh * ow
    return math.sqrt(se_t / max(wsum, 1.0)), math.sqrt(se_g / max(n_all, 1))


def train(pairs=20000, holdout=256, epochs=8, batch=32, lr=1e-3, lr_min=2e-5, warmup=200, seed=0, out=None,
          report=None, init=None, seed0=TRAIN_SEED0):
    _native.require_device()
    net = CoarseFlowNet(SHAPE, max_batch=batch, seed=seed)
    if init:
        net.set_params(CoarseFlowNet.load(init).get_params())
        log(f"resuming from {init}")
    log(f"network: {net.n_params} parameters, input {net.in_shape}, output {net.out_shape}")
    t0 = time.perf_counter()
    log(f"building {pairs} training pairs (seeds {seed0}+)")
    train_set = Dataset(net, pairs, seed0)
    log(f"building {holdout} held-out pairs (seeds {HOLDOUT_SEED0}+)")
    hold = Dataset(net, holdout, HOLDOUT_SEED0, progress_every=0)
    t_data = time.perf_counter() - t0
    log(f"data ready in {t_data:.1f} s")
    steps_per_epoch = pairs // batch
    total = steps_per_epoch * epochs
    loss = _native.DeviceBuffer((1,), np.float32)
    rng = np.random.default_rng(seed)
    hist = []
    step = 0
    t_train = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(steps_per_epoch)
        acc, nacc = 0.0, 0
        for b in order:
            i0 = int(b) * batch
            if step < warmup:
                cur = lr * (step + 1) / warmup
            else:
                cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * (step - warmup) / max(1, total - warmup)))
            net.gradients(train_set.input_ptr(i0), train_set.target_ptr(i0), train_set.weight_ptr(i0), batch,
                          loss.value)
            net.adam_step(cur)
            step += 1
            if step % 25 == 0:
                acc += float(loss.get()[0])
                nacc += 1
        _native.synchronize()
        rms_t, rms_g = evaluate(net, hold, batch)
        rec = {"epoch": ep + 1, "step": step, "train_loss": acc / max(nacc, 1), "holdout_rms_vs_plain": rms_t,
               "holdout_rms_vs_truth": rms_g, "seconds": time.perf_counter() - t_train}
        hist.append(rec)
        log(json.dumps(rec))
    t_train = time.perf_counter() - t_train
    path = out or default_weights_path(SHAPE)
    net.save(path)
    log(f"saved {path}")
    summary = {"pairs": pairs, "holdout": holdout, "epochs": epochs, "batch": batch, "lr": lr, "steps": step,
               "data_seconds": t_data, "train_seconds": t_train, "ms_per_step": 1e3 * t_train / max(step, 1),
               "history": hist}
    if report:
        with open(report, "w") as f:
            json.dump(summary, f, indent=1)
    return net, summary


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--holdout", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    ap.add_argument("--init", default=None, help="start from these weights (.npz)")
    ap.add_argument("--lr-min", type=float, default=2e-5)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--seed0", type=int, default=TRAIN_SEED0, help="first training seed")
    a = ap.parse_args(argv)
    train(a.pairs, a.holdout, a.epochs, a.batch, a.lr, a.lr_min, a.warmup, seed=a.seed, out=a.out, report=a.report,
          init=a.init, seed0=a.seed0)
    return 0


if __name__ == "__main__":
    sys.exit(main())
