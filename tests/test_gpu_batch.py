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
            choices = batch.tune(repeats=2, cache=False)
            assert choices[0] in (128, 256, 512) and batch.tuning_source == "timed"
        batch.run()
        for i in range(len(specs)):
            u, v, vi, vf = batch.result(i)
            assert np.array_equal(u, ref[i][0].astype(np.float32)), (rep, i)
            assert np.array_equal(v, ref[i][1].astype(np.float32)), (rep, i)
            assert np.array_equal(vi, ref[i][2]), (rep, i)


def test_tuning_cache_round_trip(tmp_path):
    """A timed tuning is stored under the batch's key (architecture + pair
    descriptions); a second batch of the same description applies it from
    the cache without timing, and runs to the same results; a different
    description misses; wrong choices are refused without changing the
    batch."""
    from openpiv_gpu import _native
    from openpiv_gpu import synthetic as S
    from openpiv_gpu.batch import PIVBatch
    specs = [((256, 320), 16, dict(search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2)),
             ((384, 256), 24, dict(search_size_iters=(1, 1), overlap_ratio=0.625, median_tol=2))]
    frames = []
    for i, (shape, _, _) in enumerate(specs):
        gen = S.DevicePairGenerator(shape, S.MODEL_12BIT)
        frames.append(gen.render(S.vortex_field(4300 + i, shape), 4300 + i))
    cache = tmp_path / "tune.json"
    a = PIVBatch(specs)
    a.set_frames(frames)
    ca = a.tune(repeats=1, cache=cache)
    assert a.tuning_source == "timed" and cache.exists() and len(ca) == a.choice_count()
    a.run()
    b = PIVBatch(specs)
    b.set_frames(frames)
    assert b.tuning_key() == a.tuning_key()
    cb = b.tune(repeats=1, cache=cache)
    assert b.tuning_source == "cache" and np.array_equal(ca, cb)
    b.run()
    for i in range(len(specs)):
        ra, rb = a.result(i), b.result(i)
        assert all(np.array_equal(x, y) for x, y in zip(ra, rb)), i
    c = PIVBatch(specs[:1])
    assert c.tuning_key() != a.tuning_key()
    bad = np.array(cb, np.int32)
    bad[0] = 100
    with pytest.raises(_native.NativeError):
        b.set_choices(bad)
    with pytest.raises(_native.NativeError):
        b.set_choices(cb[:-1])
