"""CPU-only checks of the C ABI library: it loads and exports every function
the public header declares (no compute call is made without a GPU)."""
import ctypes as C
import os

import pytest

from optcuts_amd import capi


def test_library_is_built_and_loads():
    assert os.path.exists(capi.LIB_PATH), "run optcuts_amd.build first"
    capi.load()


def test_every_header_function_is_exported():
    names = capi.header_functions()
    assert len(names) >= 25
    L = C.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert set(capi.header_functions()) <= set(capi._SIGS)


def test_null_context_is_rejected_without_a_gpu_call():
    L = capi.load()
    assert L.oc_ctx_destroy(None) == -1
    assert L.oc_energy(None, 0, None) == -1
    assert L.oc_last_error(None) == b"null context"


def test_status_codes_match_header():
    text = open(capi.HEADER).read()
    for code, name in capi.STATUS.items():
        if name == "OC_OK":
            assert "OC_OK = 0" in text
        else:
            assert "%s = %d" % (name, code) in text, name


def test_no_gpu_here_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.OCError):
        capi.Context(0)
