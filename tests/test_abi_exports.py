"""The C-ABI library loads and exports every symbol include/*.h declares.

No compute calls here: this runs on machines without a GPU.
"""
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(REPO, "include", h) for h in ("restmatr.h", "rest_tensors_amd.h")]
DECL = re.compile(r"^(?:void|int|int64_t|const char\s*\*)\s*\**\s*(\w+)\s*\(", re.M)


def declared_symbols():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += DECL.findall(text)
    return names


def test_headers_declare_the_seven_dropins():
    names = declared_symbols()
    for s in ("ri_ao2mo_f_", "general_dgemm_f_", "special_dgemm_f_01_", "copy_mm_",
              "copy_mr_", "copy_rm_", "copy_rr_"):
        assert s in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    import ctypes
    from rest_tensors_amd._lib import LIB_PATH, DROPIN_SIGNATURES, RT_SIGNATURES
    if not os.path.exists(LIB_PATH):
        pytest.fail(f"{LIB_PATH} not built (run python -m rest_tensors_amd.build)")
    lib = ctypes.CDLL(LIB_PATH)
    names = declared_symbols()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers every declared symbol with a signature
    from rest_tensors_amd import _lib
    bound = set(DROPIN_SIGNATURES) | set(RT_SIGNATURES) | set(_lib.EXTRA_SIGNATURES)
    assert not [n for n in names if n not in bound]


def test_upper_dim_host_function():
    from rest_tensors_amd._lib import lib
    from oracle import ri_oracle as O
    for L in list(range(0, 600)) + [524800, 524801, 131328, 8256, 2**40 + 3]:
        assert lib().rt_upper_dim(L) == O.upper_dim(L)
    assert lib().rt_upper_dim(-1) == -1
