"""The C ABI boundary, checked on the CPU (no compute calls):
  - libfsgm.so loads and exports every function include/sgm_capi.h declares
  - libStereoCuda.so exports every StereoCuda member include/StereoCuda.h declares
  - libFasterStereoCuda.so exports every FasterStereoCuda member of its header
  - the ctypes mirror of sgm_option has the header's field order and types
  - argument validation happens before the device is touched
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from fasterstereocuda_amd import _lib
from fasterstereocuda_amd.stereo import CuSGMOption, StereoCuda

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "fasterstereocuda_amd")


def _header_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sgm_\w+)\s*\(", text)))


def _exports(lib, demangle=False):
    out = subprocess.run(["nm", "-D", "--defined-only"] + (["-C"] if demangle else []) + [lib],
                         capture_output=True, text=True, check=True).stdout
    return out


def test_capi_exports_every_declared_function():
    declared = [f for f in _header_functions(os.path.join(INCLUDE, "sgm_capi.h"))]
    assert len(declared) >= 20
    syms = set(line.split()[-1] for line in _exports(os.path.join(PKG, "libfsgm.so")).splitlines())
    missing = [f for f in declared if f not in syms]
    assert not missing, f"declared but not exported: {missing}"
    lib = _lib.lib()
    for f in declared:
        assert hasattr(lib, f)
    assert set(declared) == set(_lib.SIGNATURES), "ctypes signature table out of date"


def test_stereocuda_class_exports():
    text = open(os.path.join(INCLUDE, "StereoCuda.h")).read()
    members = set(re.findall(r"\b(Init2?|Match2?|SetMinDisparity|GetDispartyRange|Release|LastError|"
                             r"GetRoiFromDispMap|get_invad_float|get_invad_short|get_level_range|"
                             r"MallocPageLockedPtr|FreePageLockedPtr)\s*\(", text))
    exported = _exports(os.path.join(PKG, "libStereoCuda.so"), demangle=True)
    for m in sorted(members):
        assert f"StereoCuda::{m}(" in exported, f"StereoCuda::{m} not exported"
    assert "StereoCuda::StereoCuda()" in exported and "StereoCuda::~StereoCuda()" in exported


def test_faster_stereo_cuda_class_exports():
    text = open(os.path.join(INCLUDE, "FasterStereoCuda.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    members = set(re.findall(r"\b(Init2?|Match2?|Release|LastError|MallocPageLockedPtr|"
                             r"FreePageLockedPtr)\s*\(", text))
    assert members == {"Init", "Init2", "Match", "Match2", "Release", "LastError",
                       "MallocPageLockedPtr", "FreePageLockedPtr"}
    exported = _exports(os.path.join(PKG, "libFasterStereoCuda.so"), demangle=True)
    for m in sorted(members):
        assert f"FasterStereoCuda::{m}(" in exported, f"FasterStereoCuda::{m} not exported"
    assert "FasterStereoCuda::FasterStereoCuda()" in exported
    # the option struct keeps the reference's field order and defaults
    body = re.search(r"typedef struct StereoOption1 \{(.*?)\} ste_opt1_t;", text, re.S).group(1)
    assert re.findall(r"(?:int|bool)\s+(\w+);", body) == [
        "width", "height", "min_disp", "max_disp", "num_layers", "do_lr_check", "do_rm_peaks",
        "do_smooth"]
    assert "num_layers(1), do_lr_check(true)" in body.replace("\n", " ").replace("  ", " ")


def test_option_struct_matches_header():
    text = open(os.path.join(INCLUDE, "sgm_capi.h")).read()
    body = re.search(r"typedef struct sgm_option \{(.*?)\} sgm_option;", text, re.S).group(1)
    fields = re.findall(r"(int32_t|float)\s+(\w+);", body)
    py = [(n, t) for n, t in _lib.SgmOption._fields_]
    assert [n for _, n in fields] == [n for n, _ in py]
    for (ct, _), (_, pt) in zip(fields, py):
        assert (ct == "float") == (pt is ctypes.c_float)
    # and the CuSGMOption mirror fills every field in order
    o = CuSGMOption.benchmark()
    c = o.to_c()
    assert (c.p1, c.p2_init, c.cs_type, c.do_lr_check, c.do_median_filter) == (10, 150, 1, 1, 1)
    assert abs(c.unique_threshold - 0.95) < 1e-7 and c.lr_threshold == 1.0


@pytest.mark.parametrize("args,code,needle", [
    ((0, 10, 0, 64), _lib.SGM_E_INVALID_ARG, "size"),
    ((100, 10, 0, 96), _lib.SGM_E_INVALID_ARG, "disp_range"),
    ((100, 10, 0, 1024), _lib.SGM_E_INVALID_ARG, "disp_range"),
    ((100, 10, 100, 64), _lib.SGM_E_INVALID_ARG, "min_disparity"),
])
def test_init_validates_before_device(args, code, needle):
    lib = _lib.lib()
    h = lib.sgm_create()
    opt = CuSGMOption.benchmark().to_c()
    rc = lib.sgm_init(h, *args, ctypes.byref(opt), 0)
    assert rc == code
    assert needle in _lib.last_error(h)
    lib.sgm_destroy(h)


def test_peak_removal_size_limit():
    lib = _lib.lib()
    h = lib.sgm_create()
    opt = CuSGMOption.benchmark(do_remove_peaks=True, peaks_ratio_threshold=0.001).to_c()
    assert lib.sgm_init(h, 50000, 50000, 0, 64, ctypes.byref(opt), 0) == _lib.SGM_E_INVALID_ARG
    assert "2^31" in _lib.last_error(h)
    lib.sgm_destroy(h)


@pytest.mark.parametrize("fields,code", [
    ({"do_remove_peaks": True, "peaks_ratio_threshold": 2.0}, _lib.SGM_E_INVALID_ARG),
    ({"do_remove_peaks": True, "peaks_ratio_threshold": -0.1}, _lib.SGM_E_INVALID_ARG),
    ({"post_filter_type": 1}, _lib.SGM_E_UNSUPPORTED),
    ({"morphology_type": 2}, _lib.SGM_E_UNSUPPORTED),
    ({"p1": -1}, _lib.SGM_E_INVALID_ARG),
    ({"p2_init": 70000}, _lib.SGM_E_INVALID_ARG),
    ({"cs_type": 3}, _lib.SGM_E_INVALID_ARG),
], ids=lambda v: "-".join(f"{k}={x}" for k, x in v.items()) if isinstance(v, dict) else str(v))
def test_option_validation(fields, code):
    lib = _lib.lib()
    h = lib.sgm_create()
    opt = CuSGMOption.benchmark(**fields).to_c()
    assert lib.sgm_init(h, 100, 50, 0, 64, ctypes.byref(opt), 0) == code
    lib.sgm_destroy(h)


def test_default_option_passes_validation():
    """A default-constructed CuSGMOption (using_constant_p2 false,
    sgm_types.h constructor) with only P1, P2 and the census type set is a
    valid option: validation passes and the call reaches the device query
    (which fails here, on a host without a GPU, with SGM_E_NO_DEVICE)."""
    lib = _lib.lib()
    h = lib.sgm_create()
    opt = CuSGMOption(p1=10, p2_init=150, cs_type=1).to_c()
    assert opt.using_constant_p2 == 0
    rc = lib.sgm_init(h, 100, 50, 0, 64, ctypes.byref(opt), 0)
    assert rc in (_lib.SGM_OK, _lib.SGM_E_NO_DEVICE, _lib.SGM_E_HIP), _lib.last_error(h)
    # disp_range 32 (LEVEL_RANGE) is accepted too: matching around an initial map
    rc = lib.sgm_init(h, 100, 50, 0, 32, ctypes.byref(opt), 0)
    assert rc in (_lib.SGM_OK, _lib.SGM_E_NO_DEVICE, _lib.SGM_E_HIP), _lib.last_error(h)
    lib.sgm_destroy(h)


def test_python_bindings_check_buffers_before_any_call():
    """Wrong shapes, dtypes, layouts or devices raise ValueError in the
    bindings, before a pointer reaches the C ABI."""
    import torch
    from fasterstereocuda_amd.batch import SgmBatch
    from fasterstereocuda_amd.pyramid import SgmPyramid
    st = StereoCuda()
    st.width, st.height = 8, 6
    ok8, okf = np.zeros((6, 8), np.uint8), np.zeros((6, 8), np.float32)
    bad = [
        (np.zeros((6, 7), np.uint8), ok8, okf, None, None),
        (ok8, ok8.astype(np.int8), okf, None, None),
        (ok8, ok8, np.zeros((6, 8), np.float64), None, None),
        (ok8, ok8, okf, np.zeros((5, 8), np.float32), None),
        (ok8, ok8, okf, None, np.zeros((6, 8), np.int32)),
        (ok8, ok8, np.zeros((8, 6), np.float32).T, None, None),
    ]
    for l, r, dl, dr, init in bad:
        with pytest.raises(ValueError):
            st.Match(l, r, dl, dr, init)
    with pytest.raises(ValueError):
        st.Match2(ok8, ok8, okf, np.zeros((6, 8), np.float32))
    cpu8, cpuf = torch.zeros((6, 8), dtype=torch.uint8), torch.zeros((6, 8))
    with pytest.raises(ValueError):
        st.match_device(cpu8, cpu8, cpuf)
    pyr = SgmPyramid.__new__(SgmPyramid)
    pyr._p, pyr.width, pyr.height = None, 8, 6
    with pytest.raises(ValueError):
        pyr.match(np.zeros((6, 9), np.uint8), ok8)
    with pytest.raises(ValueError):
        pyr.match2(ok8, np.zeros((7, 8), np.uint8))
    with pytest.raises(ValueError):
        pyr.match_device(cpu8, cpu8, cpuf)
    with pytest.raises(ValueError):
        pyr.match2_device(cpu8, cpu8, cpuf)
    with pytest.raises(ValueError):
        SgmBatch([dict(left=cpu8, right=cpu8, disp_left=cpuf, disp_range=64)],
                 CuSGMOption.benchmark())


def test_peak_strips_setting():
    lib = _lib.lib()
    assert lib.sgm_get_peak_strips() == 1
    assert lib.sgm_set_peak_strips(0) == _lib.SGM_E_INVALID_ARG
    assert lib.sgm_set_peak_strips(6) == 0 and lib.sgm_get_peak_strips() == 6
    assert lib.sgm_set_peak_strips(1) == 0


def test_window_form_setting():
    lib = _lib.lib()
    assert lib.sgm_get_window_form() == 0          # SGM_WINDOW_AUTO
    for form in (1, 2, 3, 0):
        assert lib.sgm_set_window_form(form) == 0 and lib.sgm_get_window_form() == form
    assert lib.sgm_set_window_form(4) == _lib.SGM_E_INVALID_ARG
    assert lib.sgm_set_window_form(-1) == _lib.SGM_E_INVALID_ARG
    assert lib.sgm_get_window_form() == 0


def test_wta_form_setting():
    lib = _lib.lib()
    assert lib.sgm_get_wta_form() == 0             # SGM_WTA_AUTO
    for form in (1, 2, 0):
        assert lib.sgm_set_wta_form(form) == 0 and lib.sgm_get_wta_form() == form
    assert lib.sgm_set_wta_form(3) == _lib.SGM_E_INVALID_ARG
    assert lib.sgm_set_wta_form(-1) == _lib.SGM_E_INVALID_ARG
    assert lib.sgm_get_wta_form() == 0


def test_uninitialised_handle_and_constants():
    st = StereoCuda()
    import numpy as np
    a = np.zeros((4, 4), np.uint8)
    o = np.zeros((4, 4), np.float32)
    st.width, st.height = 4, 4
    assert not st.Match(a, a, o)
    assert "not initialised" in st.last_error
    st.Release()
    st.Release()
    assert StereoCuda.get_invad_float() == -999999.0
    assert StereoCuda.get_invad_short() == -32768
    assert StereoCuda.get_level_range() == 32


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """No CPU fallback: without the HIP library every entry point raises."""
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libfsgm.so"))
    with pytest.raises(ImportError, match="no CPU fallback"):
        _lib.lib()
    with pytest.raises(ImportError):
        StereoCuda().Init(64, 32, 0, 64, CuSGMOption.benchmark())


def test_roi_from_disparity_empty_size_is_a_no_op():
    roi = _lib.SgmRoi(1, 2, 3, 4)
    d = np.zeros(4, np.float32)
    _lib.lib().sgm_roi_from_disparity(d.ctypes.data, 0, 4, ctypes.byref(roi))
    _lib.lib().sgm_roi_from_disparity(d.ctypes.data, 4, -1, ctypes.byref(roi))
    assert (roi.x, roi.y, roi.w, roi.h) == (1, 2, 3, 4)


def test_roi_from_disparity_host_scan():
    import numpy as np
    d = np.full((10, 12), -999999.0, np.float32)
    d[3, 4] = 1.0
    d[7, 9] = 2.0
    d[5, 2] = 3.0
    roi = StereoCuda.GetRoiFromDispMap(d)
    assert (roi.x, roi.y, roi.w, roi.h) == (2, 3, 8, 5)


@pytest.mark.parametrize("B,H,W,D", [(8, 375, 1242, 128), (1, 33, 40, 64), (2, 20, 64, 256)])
def test_dsgm_memory_is_bounded(B, H, W, D):
    """The differentiable SGM keeps no per-direction volume.  What the forward
    saves for the backward is two floats per pixel, the winner codes (2 bits
    per cell and direction: one byte per cell) and one argmin byte per pixel
    and direction, about a quarter of one FP32 volume; the transient
    workspace is one direction volume plus two per-pixel planes."""
    L = _lib.lib()
    npix = B * H * W
    saved = L.sgm_dsgm_checkpoint_floats(B, H, W, D)
    assert saved == -(-(npix * 8 + npix * D + npix * 4) // 4)
    assert 4 * saved < 0.3 * 4 * npix * D          # under 0.3 of a volume
    assert L.sgm_dsgm_workspace_floats(B, H, W, D) == npix * D + 2 * npix
    assert L.sgm_dsgm_workspace_floats(B, H, W, 96) == 0   # D must be 64, 128 or 256
