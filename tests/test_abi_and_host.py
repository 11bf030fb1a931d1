"""CPU-side checks: the C-ABI library loads and exports every declared
symbol, the ctypes structs match the C layout, and the host-side settings
logic mirrors the reference interface (no compute calls)."""

import os
import re
import subprocess

import numpy as np
import pytest

from openpiv_gpu import _native
from openpiv_gpu import gpu_process as gp
from oracle import piv_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "openpiv_hip.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(piv_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_match_binding():
    assert _declared_functions() == sorted(_native.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("HIP library not built")
    lib = _native.load_library()
    for name in _declared_functions():
        assert hasattr(lib, name), name


STRUCTS = (("piv_pass_desc", "PassDesc"), ("piv_plan_desc", "PlanDesc"), ("piv_synth_field", "SynthField"),
           ("piv_synth_model", "SynthModel"))


def test_struct_layout_matches_c(tmp_path):
    """sizeof and every field offset of each C struct equal the ctypes mirror."""
    import ctypes
    lines, expect = [], []
    for cname, pyname in STRUCTS:
        cls = getattr(_native, pyname)
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append(ctypes.sizeof(cls))
        for fname, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
            expect.append(getattr(cls, fname).offset)
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "openpiv_hip.h"\nint main(void){'
                   + "".join(lines) + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = [int(t) for t in subprocess.check_output([str(exe)]).split()]
    assert got == expect


def test_no_device_means_loud_failure():
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("HIP library not built")
    if _native.device_count() > 0:
        pytest.skip("a GPU is visible")
    piv = gp.PIVGPU((128, 128), 32, dtype_f="float32")
    with pytest.raises(_native.NativeError):
        piv.run_host(np.zeros((128, 128), np.uint8), np.zeros((128, 128), np.uint8))


def test_plan_desc_geometry_matches_oracle():
    piv = gp.PIVGPU((2048, 2048), 16, search_size_iters=(1, 1, 1), dtype_f="float32")
    d = piv._plan_desc(_native.PIV_U16)
    assert d.n_passes == 3
    for k, ss in enumerate((64, 32, 16)):
        g = O.field_geometry((2048, 2048), ss, 0.5)
        p = d.passes[k]
        assert (p.search_size, p.spacing, p.x_offset, p.y_offset) == (ss, g.spacing, *g.offset)
        assert (p.field_m, p.field_n) == g.field_shape
        assert p.fft_size == O.fft_size(ss)
        assert p.smooth == 1 and p.validate == 1 and p.num_replacing_iters == 1
    x, y = piv.coords
    np.testing.assert_array_equal(x.get(), O.field_geometry((2048, 2048), 16, 0.5).x)


def test_per_size_settings_follow_size_groups():
    piv = gp.PIVGPU((512, 512), 16, search_size_iters=(1, 2), n_fft=(2, 3), median_tol=(None, 1.5),
                    deforming_par=(0.5, 1.0), dtype_f="float32")
    d = piv._plan_desc(_native.PIV_U8)
    assert [d.passes[k].search_size for k in range(3)] == [32, 16, 16]
    assert [d.passes[k].fft_size for k in range(3)] == [64, 64, 64]
    assert [d.passes[k].validate for k in range(3)] == [0, 1, 1]
    assert d.passes[2].median_tol == pytest.approx(1.5)
    assert d.passes[1].deforming_par == 1.0


@pytest.mark.parametrize("kwargs,msg", [
    (dict(search_size_iters=(0,)), "search_size_iters must be a tuple of positive int numbers."),
    (dict(overlap_ratio=1.5), "overlap_ratio must be a tuple of real values between 0 and 1, defined for all passes."),
    (dict(subpixel_method="bogus"), "subpixel_method for every pass must be one of"),
    (dict(dt=-1.0), "dt must be a real number greater than 0."),
    (dict(center=1), "center must have a bool value."),
])
def test_argument_assertions_match_reference_messages(kwargs, msg):
    with pytest.raises(AssertionError, match=re.escape(msg)):
        gp.piv_gpu((256, 256), 16, **kwargs)


def test_min_search_size_and_frame_shape_assertions():
    with pytest.raises(AssertionError, match="min_search_size must be an int number greater than 8."):
        gp.piv_gpu((256, 256), 4)
    with pytest.raises(AssertionError, match="frame_shape must be a tuple of int values greater than max_search_size"):
        gp.piv_gpu((32, 256), 16, search_size_iters=(1, 1, 1))


@pytest.mark.parametrize("kwargs,name", [
    (dict(validation_size=9), "validation_size"),
    (dict(replacing_size=9), "replacing_size"),
])
def test_unimplemented_options_raise(kwargs, name):
    with pytest.raises(NotImplementedError, match=name):
        gp.piv_gpu((256, 256), 16, **kwargs)


def test_kernel_sizes_reach_the_plan_description():
    """validation_size / replacing_size per pass (gpu_process.py:182-187,
    239-244) land in the pass descriptors; a half-width above a field
    dimension fails the reference's assertion (gpu_validation.py:74, 273)."""
    piv = gp.PIVGPU((256, 256), 16, search_size_iters=(1, 1), validation_size=(3, 2), replacing_size=(1, 4))
    d = piv._plan_desc(_native.PIV_U8)
    assert [(d.passes[k].validation_size, d.passes[k].replacing_size) for k in range(2)] == [(3, 1), (2, 4)]
    with pytest.raises(AssertionError, match="size cannot exceed field_shape"):
        gp.PIVGPU((64, 256), 32, search_size_iters=(1,), validation_size=4)._plan_desc(_native.PIV_U8)


def test_fp64_settings_on_the_host():
    """dtype_f other than the string "float32" selects FP64 processing
    (gpu_process.py:552): double coordinates, an FP64 plan description with
    the exact scaling factors, float64 frames accepted raw."""
    piv = gp.PIVGPU((256, 256), 16, search_size_iters=(1, 1), dtype_f="float64", dt=0.3, scaling_par=2.0)
    assert piv.fp64 and piv.dtype_f == np.float64
    assert piv.piv_fields[-1].x.dtype == np.float64
    d = piv._plan_desc(_native.PIV_F64)
    assert d.precision == _native.PIV_FP64 and d.scaling_par == 2.0 and d.dt == 0.3
    assert piv._prepare_frame(np.zeros((256, 256), np.float64)).dtype == np.float64
    assert piv._prepare_frame(np.zeros((256, 256), np.int32)).dtype == np.float64
    assert gp.PIVGPU((256, 256), 16, dtype_f=np.float32).fp64        # not the string: FP64, as the reference
    p32 = gp.PIVGPU((256, 256), 16, dtype_f="float32")
    assert not p32.fp64 and p32._prepare_frame(np.zeros((256, 256), np.float64)).dtype == np.float32
    assert p32._plan_desc(_native.PIV_U8).precision == _native.PIV_FP32
    with pytest.raises(NotImplementedError, match="predictor"):
        gp.PIVGPU((256, 256), 16, dtype_f="float64", predictor=object())


def test_mask_settings_on_the_host():
    """The frame mask reaches every pass's field (gpu_process.py:949-951) and
    the class rejects what the plan cannot take."""
    mask = np.zeros((256, 256), bool)
    mask[100:140, 60:90] = True
    piv = gp.piv_gpu((256, 256), 16, search_size_iters=(1, 2), mask=mask.astype(np.uint8))
    for f in piv.gpu_process.piv_fields:
        fm = f.field_mask
        assert fm.shape == f.field_shape and fm.dtype == bool
        assert np.array_equal(fm, mask[f.y.astype(int), f.x.astype(int)])
    assert np.array_equal(piv.field_mask, piv.gpu_process.field_mask)
    with pytest.raises(ValueError, match="mask has shape"):
        gp.PIVGPU((256, 256), 16, mask=np.zeros((128, 256), bool))
    with pytest.raises(NotImplementedError, match="predictor"):
        gp.PIVGPU((256, 256), 16, mask=mask, predictor=object())


def test_plan_desc_carries_every_validation_setting():
    piv = gp.PIVGPU((256, 256), 16, search_size_iters=(1, 2), s2n_method="peak2mean", s2n_size=3, s2n_tol=1.3,
                    median_tol=None, mad_tol=2.5, mean_tol=(1.0, None), rms_tol=3.0, replacing_method="median",
                    num_replacing_iters=3, revalidate=True, dtype_f="float32")
    d = piv._plan_desc(_native.PIV_U8)
    assert d.n_passes == 3
    for k, p in enumerate(d.passes[:3]):
        assert p.validate == 1 and np.isnan(p.median_tol)
        assert p.s2n_method == 1 and p.s2n_size == 3 and np.isclose(p.s2n_tol, 1.3) and np.isclose(p.mad_tol, 2.5)
        assert np.isclose(p.rms_tol, 3.0) and p.replacing_method == 1 and p.num_replacing_iters == 3
        assert p.revalidate == 1
    # per size group (gpu_process.py:702-711): passes 1 and 2 share the second entry
    assert np.isclose(d.passes[0].mean_tol, 1.0) and np.isnan(d.passes[1].mean_tol) and np.isnan(d.passes[2].mean_tol)
    none = gp.PIVGPU((256, 256), 16, median_tol=None, dtype_f="float32")._plan_desc(_native.PIV_U8)
    assert none.passes[0].validate == 0


def test_shrink_ratio_sets_first_pass_window():
    piv = gp.PIVGPU((256, 256), 16, search_size_iters=(1, 1), shrink_ratio=0.6, dtype_f="float32")
    f0, f1 = piv.piv_fields
    assert (f0.search_size, f0.window_size) == (32, 20) and (f1.search_size, f1.window_size) == (16, 16)
    d = piv._plan_desc(_native.PIV_U8)
    assert d.passes[0].window_size == 20 and d.passes[1].window_size == 0
    # int(64 * 0.1) = 6 -> clipped to 8; odd sizes round up
    assert gp.PIVGPU((256, 256), 64, shrink_ratio=0.1, dtype_f="float32").piv_fields[0].window_size == 8
    assert gp.PIVGPU((256, 256), 32, shrink_ratio=0.55, dtype_f="float32").piv_fields[0].window_size == 18


def test_gpu_piv_alias_and_frame_assertion():
    assert gp.gpu_piv is gp.piv_gpu
    piv = gp.piv_gpu((128, 128), 32, dtype_f="float32")
    with pytest.raises(AssertionError, match="Both frames must be an ndarray of real values with shape"):
        piv(np.zeros((64, 64)), np.zeros((64, 64)))
    assert piv.field_shape == (7, 7)
    assert not piv.val_locations.any() and not piv.outliers.any()
    assert piv.field_mask.shape == (7, 7) and not piv.field_mask.any()


def test_batch_tuning_cache_helpers(tmp_path, monkeypatch):
    """The tuning cache's path rules, JSON round trip (atomic replace) and the
    stable description of pair keyword arguments."""
    from openpiv_gpu import batch as B
    monkeypatch.setenv("OPENPIV_HIP_TUNE_CACHE", str(tmp_path / "x.json"))
    assert B._cache_path(True) == str(tmp_path / "x.json")
    assert B._cache_path(False) is None and B._cache_path(tmp_path / "y.json") == str(tmp_path / "y.json")
    p = str(tmp_path / "sub" / "c.json")
    assert B._read_cache(p) == {}
    B._write_cache(p, {"k": [256, 1]})
    assert B._read_cache(p) == {"k": [256, 1]}
    (tmp_path / "bad.json").write_text("not json")
    assert B._read_cache(str(tmp_path / "bad.json")) == {}
    kw = dict(search_size_iters=(1, 1), overlap_ratio=0.5, mask=np.ones((4, 4), bool))
    j1, j2 = B._jsonable(kw), B._jsonable(dict(reversed(list(kw.items()))))
    assert j1 == j2 and isinstance(j1["mask"], str) and j1["search_size_iters"] == [1, 1]
