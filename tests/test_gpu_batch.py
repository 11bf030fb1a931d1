"""Batches of uneven pairs (one deformation and one correlation launch per
pass level over all pairs) against each pair run on its own plan: results
must be identical, before and after the launch-shape autotuner, eager and
graph-replayed."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batch_matches_single_pair_runs():
    from openpiv_gpu import synthetic as S
    from openpiv_gpu.batch import PIVBatch
    from openpiv_gpu.gpu_process import PIVGPU
    specs = [((256, 320), 16, dict(search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2)),
             ((384, 256), 24, dict(search_size_iters=(1, 1), overlap_ratio=0.625, median_tol=2)),
             ((512, 512), 16, dict(search_size_iters=(1, 1, 1), overlap_ratio=0.7, median_tol=2,
                                   replacing_method="median", num_replacing_iters=2)),
             ((320, 448), 16, dict(search_size_iters=(1, 1), overlap_ratio=0.5, median_tol=2, s2n_tol=1.1,
                                   s2n_method="peak2mean")),
             ((256, 256), 24, dict(search_size_iters=(1, 1), overlap_ratio=0.5, median_tol=2))]
    frames = []
    for i, (shape, _, _) in enumerate(specs):
        gen = S.DevicePairGenerator(shape, S.MODEL_12BIT)
        frames.append(gen.render(S.vortex_field(4200 + i, shape), 4200 + i))
    batch = PIVBatch(specs)
    batch.set_frames(frames)
    ref = []
    for (shape, mss, kw), (fa, fb) in zip(specs, frames):
        piv = PIVGPU(shape, mss, dtype_f="float32", **kw)
        u, v = piv.run_host(fa.get(), fb.get())
        ref.append((u, v, np.asarray(piv.init_val_locations.get() if hasattr(piv.init_val_locations, "get")
                                     else piv.init_val_locations, bool)))
    for rep in range(3):   # eager, capture + replay, replay
        if rep == 1:
            choices = batch.tune(repeats=2)
            assert choices[0] in (128, 256, 512)
        batch.run()
        for i in range(len(specs)):
            u, v, vi, vf = batch.result(i)
            assert np.array_equal(u, ref[i][0].astype(np.float32)), (rep, i)
            assert np.array_equal(v, ref[i][1].astype(np.float32)), (rep, i)
            assert np.array_equal(vi, ref[i][2]), (rep, i)
