"""Native baseline-TIFF decoding (the frame input of pair streams) against
PIL on files PIL writes: 8/16-bit unsigned, 32-bit float, both byte orders,
single and multiple strips; rejection of what the decoder does not cover."""

import os

import numpy as np
import pytest

from openpiv_gpu import _native

pytestmark = pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason="HIP library not built")
Image = pytest.importorskip("PIL.Image")


def _write(path, arr, mode=None, **kw):
    Image.fromarray(arr, mode=mode).save(path, **kw) if mode else Image.fromarray(arr).save(path, **kw)


@pytest.mark.parametrize("dtype,mode", [(np.uint8, None), (np.uint16, "I;16"), (np.uint16, "I;16B"),
                                        (np.float32, "F")])
@pytest.mark.parametrize("shape", [(37, 53), (512, 256)])
def test_read_tiff_matches_pil(tmp_path, dtype, mode, shape):
    from openpiv_gpu.io import read_tiff, tiff_info
    rng = np.random.default_rng(1)
    if dtype == np.float32:
        a = rng.standard_normal(shape).astype(np.float32)
    else:
        a = rng.integers(0, np.iinfo(dtype).max, shape).astype(dtype)
    p = str(tmp_path / "f.tif")
    if mode == "I;16B":
        Image.frombytes("I;16B", shape[::-1], a.astype(">u2").tobytes()).save(p)
    elif mode:
        Image.frombytes(mode, shape[::-1], a.tobytes()).save(p)
    else:
        Image.fromarray(a).save(p)
    ref = np.asarray(Image.open(p))
    h, w, dt = tiff_info(p)
    assert (h, w) == shape and dt == np.dtype(dtype)
    got = read_tiff(p)
    assert got.dtype == np.dtype(dtype) and np.array_equal(got, ref.astype(dtype)) and np.array_equal(got, a)


def test_read_tiff_multiple_strips(tmp_path):
    from openpiv_gpu.io import read_tiff
    a = np.arange(300 * 200, dtype=np.uint16).reshape(300, 200)
    p = str(tmp_path / "s.tif")
    # libtiff splits a 16-bit image into strips of ~8 KB unless told otherwise
    Image.frombytes("I;16", (200, 300), a.tobytes()).save(p, tiffinfo={278: 7})
    assert np.array_equal(read_tiff(p), a)


def test_read_tiff_rejects_compressed_and_garbage(tmp_path):
    from openpiv_gpu.io import read_tiff
    a = np.zeros((16, 16), np.uint8)
    p = str(tmp_path / "c.tif")
    Image.fromarray(a).save(p, compression="tiff_lzw")
    with pytest.raises(_native.NativeError, match="compression"):
        read_tiff(p)
    q = tmp_path / "g.tif"
    q.write_bytes(b"not a tiff at all")
    with pytest.raises(_native.NativeError, match="byte-order"):
        read_tiff(str(q))
    with pytest.raises(_native.NativeError, match="cannot open"):
        read_tiff(str(tmp_path / "missing.tif"))
