"""Pair streams from TIFF files (overlapped decode / upload / compute) give
exactly the per-pair results of the host path."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
Image = pytest.importorskip("PIL.Image")


def test_pair_stream_matches_host_runs(tmp_path):
    from openpiv_gpu import synthetic as S
    from openpiv_gpu.gpu_process import PIVGPU
    from openpiv_gpu.io import PairStream
    shape = (512, 512)
    gen = S.DevicePairGenerator(shape, S.MODEL_12BIT)
    pairs, frames = [], []
    for i in range(7):
        fa, fb = gen.render(S.vortex_field(4300 + i, shape), 4300 + i)
        a, b = fa.get(), fb.get()
        pa, pb = str(tmp_path / f"{i}_a.tif"), str(tmp_path / f"{i}_b.tif")
        Image.frombytes("I;16", shape[::-1], a.tobytes()).save(pa)
        Image.frombytes("I;16", shape[::-1], b.tobytes()).save(pb)
        pairs.append((pa, pb))
        frames.append((a, b))
    kw = dict(search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float32")
    ref = PIVGPU(shape, 16, **kw)
    stream = PairStream(PIVGPU(shape, 16, **kw), depth=3, workers=2)
    seen = 0
    for i, u, v, flags in stream.run(pairs):
        ru, rv = ref.run_host(*frames[i])
        assert np.array_equal(u, ru.astype(np.float32)) and np.array_equal(v, rv.astype(np.float32)), i
        assert np.array_equal(flags, np.asarray(ref.init_val_locations, bool))
        seen += 1
    assert seen == len(pairs)
