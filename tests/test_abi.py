"""CPU checks of the C ABI boundary: the library loads and exports every entry
point include/pascal_poisson_fft.h declares (no compute without a GPU)."""
import ctypes
import os
import re

import pytest

from pascal_poisson_fft_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pascal_poisson_fft.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ppf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("ppf_plan_create", "ppf_solve", "ppf_plan_destroy", "ppf_tdma_many_solve", "ppf_malloc"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(api.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(api.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(api.LIB_PATH), reason="library not built")
def test_error_strings():
    assert api.error_string(0) == "success"
    assert "extent" in api.error_string(-2)
    assert api.error_string(12345) == "unknown error"


# ---- failure handling through the C ABI, without a GPU ---------------------
# Every entry point checks its arguments before its first HIP call, so these
# run on the CPU.  The reference prints and stops on the same inputs (odd n:
# 01_Fortran/examples/global.f90:114-120; plan sizes: pascal_tdma_cuda.f90:
# 168-190); here each returns its code.

needs_lib = pytest.mark.skipif(not os.path.exists(api.LIB_PATH), reason="library not built")


@needs_lib
@pytest.mark.parametrize("n1,n2,n3,code", [
    (33, 32, 32, api.PPF_ERR_SIZE),     # odd x
    (32, 22, 32, api.PPF_ERR_SIZE),     # factor 11: no radix for it
    (1280, 32, 32, api.PPF_ERR_SIZE),   # above 1024
    (4, 32, 32, api.PPF_ERR_SIZE),      # below 8
    (32, 32, 1, api.PPF_ERR_SIZE),      # z below 2
    (32, 32, 7, api.PPF_ERR_SIZE),      # odd z
    (32, 32, 4098, api.PPF_ERR_SIZE),   # z above 4096
    (0, 32, 32, api.PPF_ERR_INVALID_ARG),
    (32, -8, 32, api.PPF_ERR_INVALID_ARG),
])
def test_plan_rejects_extents(n1, n2, n3, code):
    with pytest.raises(api.PPFError) as e:
        api.FFTPoisson(n1, n2, n3)
    assert e.value.code == code


@needs_lib
def test_plan_rejects_lengths_and_precision():
    for L in ((0.0, 1.0, 1.0), (1.0, -1.0, 1.0), (1.0, 1.0, float("nan"))):
        with pytest.raises(api.PPFError) as e:
            api.FFTPoisson(32, 32, 32, L)
        assert e.value.code == api.PPF_ERR_INVALID_ARG
    with pytest.raises(api.PPFError) as e:
        api.FFTPoisson(32, 32, 32, precision=7)
    assert e.value.code == api.PPF_ERR_INVALID_ARG


@needs_lib
def test_entry_points_reject_null_arguments():
    L = api.lib()
    h = ctypes.c_void_p()
    assert L.ppf_plan_create(None, ctypes.byref(h)) == api.PPF_ERR_INVALID_ARG
    d = api.Desc(32, 32, 32, 1.0, 1.0, 1.0, 0, 0, 0, api.PPF_PREC_FP64, 0.0, 0)
    assert L.ppf_plan_create(ctypes.byref(d), None) == api.PPF_ERR_INVALID_ARG
    st = api.Stats()
    assert L.ppf_solve(None, None, None, None, None, None, None, None, ctypes.byref(st)) == api.PPF_ERR_INVALID_ARG
    r = ctypes.c_double()
    assert L.ppf_residual(None, None, None, None, None, None, ctypes.byref(r)) == api.PPF_ERR_INVALID_ARG
    assert L.ppf_solve_noalloc(ctypes.byref(d), None, None, None, None, None, None, None) == api.PPF_ERR_INVALID_ARG
    assert L.ppf_plan_destroy(None) == api.PPF_ERR_INVALID_ARG


@needs_lib
def test_batch_and_thomas_plans_reject_empty_sizes():
    L = api.lib()
    h = ctypes.c_void_p()
    assert L.ppf_batch_create(0, None, None, 0, ctypes.byref(h)) == api.PPF_ERR_INVALID_ARG
    assert L.ppf_batch_create(2, None, None, 0, ctypes.byref(h)) == api.PPF_ERR_INVALID_ARG
    assert L.ppf_tdma_plan_many_create(0, 8, 8, ctypes.byref(h)) == api.PPF_ERR_INVALID_ARG
    assert L.ppf_tdma_plan_many_create(8, 8, 0, ctypes.byref(h)) == api.PPF_ERR_INVALID_ARG
