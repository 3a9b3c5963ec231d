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


def test_merged_modes_reproduce_view_offsets():
    """The fused mode group of a strided view enumerates, in row-major
    order, exactly the element offsets of the view (host logic of the
    permute launches; no GPU call)."""
    import itertools
    import numpy as np
    from tncontract_amd._lib import _merged_modes
    rng = np.random.default_rng(0)
    for trial in range(200):
        nd = int(rng.integers(1, 11))
        shape = [int(x) for x in rng.integers(1, 4, nd)]
        base = np.arange(int(np.prod(shape))).reshape(shape)
        perm = list(rng.permutation(nd))
        view = base.transpose(perm)
        if trial % 3 == 0 and view.ndim and view.shape[0] > 1:
            view = view[1:]
        strides = [s // base.itemsize for s in view.strides]
        dims, st = _merged_modes(view.shape, strides)
        assert all(d > 1 for d in dims)
        off0 = view.reshape(-1)[0] if view.size else 0
        offs = [off0 + sum(i * s for i, s in zip(idx, st))
                for idx in itertools.product(*[range(d) for d in dims])]
        assert offs == list(view.reshape(-1))
