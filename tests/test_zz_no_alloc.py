"""piv_stage_replace_noalloc (the replacement loop on the library's fixed
device workspace, no allocation) against piv_stage_replace (the same loop on
buffers it allocates): identical results on two calls in a row, the second
with a smaller field, for every replacement method with and without
revalidation; sizes above the workspace are rejected."""

import ctypes

import numpy as np
import pytest

from openpiv_gpu import _native

pytestmark = pytest.mark.gpu
gu = pytest.importorskip("gpu_util")


def _field(seed, shape, n_out):
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    u = (2 * np.sin(xx / 7.0) + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    v = (1.5 * np.cos(yy / 5.0) + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    idx = rng.choice(u.size, n_out, replace=False)
    u.flat[idx] += rng.choice([-1, 1], n_out) * rng.uniform(3, 9, n_out)
    v.flat[idx[::2]] += rng.uniform(-8, 8, idx[::2].size)
    return u, v


def _run(entry, u, v, flags, desc, stream=None):
    lib = _native.load_library()
    m, n = u.shape
    du, dv, df = gu.dev(u), gu.dev(v), gu.dev(flags.astype(np.uint8))
    fn = getattr(lib, entry)
    _native.check(fn(gu.vp(du), gu.vp(dv), gu.vp(df), m, n, ctypes.byref(desc), None,
                     None if stream is None else stream.handle))
    if stream is not None:
        stream.synchronize()
    _native.check(lib.piv_synchronize())
    return du.get(), dv.get(), df.get()


@pytest.mark.parametrize("method,n_iters,revalidate", [("spring", 1, False), ("spring", 3, True),
                                                       ("median", 2, True), ("mean", 3, False)])
def test_noalloc_replace_equals_allocating_replace(method, n_iters, revalidate):
    from oracle import piv_oracle as O
    desc = gu.tests_desc(median_tol=1.0, method=method, n_iters=n_iters, revalidate=revalidate)
    stream = _native.Stream()
    # 255 x 255 is the largest field of the 2048^2, 16 px pass; then a smaller one
    for call, (shape, n_out) in enumerate((((255, 255), 2500), ((63, 71), 150))):
        u, v = _field(7 + call, shape, n_out)
        flags = O.validate(u, v, median_tol=1.0)
        assert flags.any()
        want = _run("piv_stage_replace", u, v, flags, desc)
        got = _run("piv_stage_replace_noalloc", u, v, flags, desc, stream)
        for w, g, name in zip(want, got, ("u", "v", "flags")):
            assert np.array_equal(w, g), (call, name)


def test_noalloc_rejects_fields_beyond_its_workspace():
    lib = _native.load_library()
    d = gu.tests_desc()
    buf = _native.DeviceBuffer((16,), np.float32)
    r = lib.piv_stage_replace_noalloc(buf.value, buf.value, buf.value, 257, 256, ctypes.byref(d), None, None)
    assert r == -1 and b"65536" in lib.piv_last_error()
