"""CPU checks of the C-ABI boundary: the library loads and exports every
function include/tncontract_amd.h declares (no compute calls: no GPU here)."""
import ctypes
import os
import re

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tncontract_amd.h")
LIB = os.path.join(ROOT, "tncontract_amd", "libtncontract_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(tn_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ["tn_contract", "tn_qr", "tn_svd", "tn_permute", "tn_sumsq", "tn_last_error"]:
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (bind to torch's HIP runtime first)
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_error_path_without_gpu_calls():
    import torch  # noqa: F401
    from tncontract_amd import _lib as L
    lib = L.lib()
    assert lib.tn_version() >= 1
    # invalid modes are rejected on the host before any device work
    m = L.tn_modes()
    m.nmodes = 99
    rc = lib.tn_permute(None, ctypes.byref(m), None, 0, None)
    assert rc == L.TN_ERR_INVALID
    assert b"modes" in lib.tn_last_error()
    with pytest.raises(ValueError):
        L.check(rc)
