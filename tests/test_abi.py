"""The C ABI library loads, exports every entry point include/*.h declares,
and validates arguments on the host before touching the GPU."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="cgsense.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cgs_[A-Za-z_0-9]+)\s*\(", src)))


HEADERS = sorted(h for h in os.listdir(os.path.join(ROOT, "include")) if h.endswith(".h"))


def test_header_declares_expected():
    names = declared_functions()
    for n in ["cgs_plan_create", "cgs_set_trajectory", "cgs_E", "cgs_EH", "cgs_solve",
              "cgs_nfft_forward", "cgs_nfft_adjoint", "cgs_nfft_inverse_ref"]:
        assert n in names


@pytest.mark.parametrize("header", HEADERS)
def test_library_exports_all_symbols(cgs, header):
    L = cgs.load_library()
    names = declared_functions(header)
    assert names
    for name in names:
        assert hasattr(L, name), name


def test_python_mirror_lists_all_symbols():
    from codeare_amd.cgsense import EXPORTED
    from codeare_amd.warmstart import UNET_EXPORTED
    assert sorted(EXPORTED) == declared_functions("cgsense.h")
    assert sorted(UNET_EXPORTED) == declared_functions("cgs_unet.h")


def test_unet_argument_validation_without_gpu(cgs):
    from codeare_amd.warmstart import _lib
    L = _lib()
    h = ct.c_void_p()
    # 4 levels need sides that are multiples of 128
    assert L.cgs_unet_create(ct.byref(h), 96, 96, 4, 32, 2, 1) == 5
    # only 32 base channels are built
    assert L.cgs_unet_create(ct.byref(h), 128, 128, 4, 16, 2, 1) == 5
    assert L.cgs_unet_create(ct.byref(h), 128, 128, 1, 32, 2, 1) == 5
    assert L.cgs_unet_create(None, 128, 128, 4, 32, 2, 1) == 1
    assert L.cgs_unet_destroy(None) == 0


def test_argument_validation_without_gpu(cgs):
    L = cgs.load_library()
    h = ct.c_void_p()
    N = (ct.c_int * 4)(16, 16, 16, 16)
    # d outside {2, 3}
    assert L.cgs_plan_create(ct.byref(h), 4, ct.cast(N, ct.c_void_p), 100, 2.0, 3, 1, 1, None, 0) == 5
    # sigma below 1
    assert L.cgs_plan_create(ct.byref(h), 2, ct.cast(N, ct.c_void_p), 100, 0.5, 3, 1, 1, None, 0) == 1
    # kernel width outside 1..4
    assert L.cgs_plan_create(ct.byref(h), 2, ct.cast(N, ct.c_void_p), 100, 2.0, 7, 1, 1, None, 0) == 5
    # odd image side
    N2 = (ct.c_int * 2)(15, 16)
    assert L.cgs_plan_create(ct.byref(h), 2, ct.cast(N2, ct.c_void_p), 100, 2.0, 3, 1, 1, None, 0) == 1
    # oversampled size with a factor 5
    N3 = (ct.c_int * 2)(20, 20)
    assert L.cgs_plan_create(ct.byref(h), 2, ct.cast(N3, ct.c_void_p), 100, 2.0, 3, 1, 1, None, 0) == 5
    # several frames without offsets
    assert L.cgs_plan_create(ct.byref(h), 2, ct.cast(N, ct.c_void_p), 100, 2.0, 3, 1, 3, None, 0) == 1
    # null plan pointer
    assert L.cgs_plan_create(None, 2, ct.cast(N, ct.c_void_p), 100, 2.0, 3, 1, 1, None, 0) == 1
    assert L.cgs_status_string(5).decode() == "unsupported parameters"
    assert len(L.cgs_last_error().decode()) > 0
    # destroy of a null plan is a no-op
    assert L.cgs_plan_destroy(None) == 0
