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
