"""CPU-side checks of the boundary: the C ABI library loads, exports every
function include/tq_hip.h declares, rejects bad arguments before touching a
device, and the product package never reaches the oracle."""
import ctypes
import os
import re

import pytest

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HEADER = os.path.join(REPO, "include", "tq_hip.h")
PKG = os.path.join(REPO, "turboquant-space_amd")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(tq_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from turboquant import capi
    L = capi.lib()
    names = declared()
    assert len(names) >= 12
    for name in names:
        assert hasattr(L, name), name
    assert sorted(capi.EXPORTED) == names


def test_bad_arguments_fail_before_device_use():
    from turboquant import capi
    L = capi.lib()
    h = ctypes.c_void_p()
    assert L.tq_space_create(64, 1, 42, 137, ctypes.byref(h)) == capi.TQ_EINVAL
    assert b"bits_per_coord" in L.tq_last_error()
    assert L.tq_space_create(64, 10, 42, 137, ctypes.byref(h)) == capi.TQ_EINVAL
    assert L.tq_space_create(0, 4, 42, 137, ctypes.byref(h)) == capi.TQ_EINVAL
    assert L.tq_space_create(8193, 4, 42, 137, ctypes.byref(h)) == capi.TQ_EUNSUPPORTED
    assert L.tq_space_create(64, 4, 42, 137, None) == capi.TQ_EINVAL
    assert L.tq_encode_batch(None, None, 1, None, 0, None) == capi.TQ_EINVAL
    assert L.tq_device_count() >= 0


def test_dropin_module_imports_and_validates_bits():
    import turboquant
    assert hasattr(turboquant, "TurboQuantSpace")
    with pytest.raises(ValueError, match="bits_per_coord"):
        turboquant.TurboQuantSpace(64, bits_per_coord=1)


def test_product_never_touches_the_oracle():
    for root, _, files in os.walk(PKG):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(root, f)).read()
                assert "oracle" not in text.lower(), os.path.join(root, f)


def test_search_entry_points_validate_before_device_use():
    from turboquant import capi
    L = capi.lib()
    h = ctypes.c_void_p()
    assert L.tq_search_plan_create(None, 0, 10, capi.TQ_MODE_MFMA, None,
                                   ctypes.byref(h)) == capi.TQ_EINVAL
    assert b"no jobs" in L.tq_last_error()
    jobs = capi.make_jobs([(0, 0, 1, 0, 1, 0, 0)])
    assert L.tq_search_plan_create(jobs, 1, 10, capi.TQ_MODE_MFMA, None,
                                   ctypes.byref(h)) == capi.TQ_EINVAL
    assert b"no space" in L.tq_last_error()
    assert L.tq_search_plan_create(jobs, 1, 0, capi.TQ_MODE_MFMA, None,
                                   ctypes.byref(h)) == capi.TQ_EINVAL
    assert L.tq_search_plan_create(jobs, 1, 10, 7, None, ctypes.byref(h)) == capi.TQ_EINVAL
    bad = capi.SearchTuning(48, 0)
    assert L.tq_search_plan_create(jobs, 1, 10, capi.TQ_MODE_MFMA, ctypes.byref(bad),
                                   ctypes.byref(h)) == capi.TQ_EINVAL
    bad_form = capi.SearchTuning(0, 0, 3)
    assert L.tq_search_plan_create(jobs, 1, 10, capi.TQ_MODE_MFMA, ctypes.byref(bad_form),
                                   ctypes.byref(h)) == capi.TQ_EINVAL
    assert b"form must be" in L.tq_last_error()
    assert L.tq_topk_rows(None, 1, 1, 129, None, None, None) == capi.TQ_EINVAL
    assert L.tq_topk_rows(None, 70000, 1, 10, None, None, None) == capi.TQ_EINVAL
    assert L.tq_search_plan_run(None, None) == capi.TQ_EINVAL
    assert L.tq_search_plan_destroy(None) == capi.TQ_OK
