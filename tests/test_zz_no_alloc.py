"""tn_svd_noalloc (the allocation-free SVD entry point) against tn_svd on the
same inputs: two calls in a row on the one scratch array, the second with a
smaller matrix, each equal to tn_svd's result to FP64 rounding; sizes above
the scratch are refused before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from tncontract_amd import _lib as L
    return L


def _run(L, a, noalloc):
    from tncontract_amd import tensor as tsr
    d = torch.from_numpy(a).cuda()
    rows, cols, m, n = tsr._matrix_modes(d, 1)
    k = min(m, n)
    u = torch.empty((m, k), dtype=torch.float64, device=d.device)
    s = torch.empty((k,), dtype=torch.float64, device=d.device)
    vt = torch.empty((k, n), dtype=torch.float64, device=d.device)
    sweeps = ctypes.c_int(0)
    lib = L.lib()
    if noalloc:
        rc = lib.tn_svd_noalloc(L.ptr(d), ctypes.byref(rows), ctypes.byref(cols), m, n, L.ptr(u),
                                L.ptr(s), L.ptr(vt), L.TN_F64, 0.0, 0, ctypes.byref(sweeps),
                                L.stream_ptr())
    else:
        ws = L.workspace(lib.tn_svd_workspace(m, n, L.TN_F64))
        rc = lib.tn_svd(L.ptr(d), ctypes.byref(rows), ctypes.byref(cols), m, n, L.ptr(u), L.ptr(s),
                        L.ptr(vt), L.TN_F64, 0.0, 0, ctypes.byref(sweeps), L.ptr(ws), ws.numel(),
                        L.stream_ptr())
    L.check(rc)
    torch.cuda.synchronize()
    return u.cpu().numpy(), s.cpu().numpy(), vt.cpu().numpy(), sweeps.value


def test_noalloc_svd_equals_tn_svd_twice(lib):
    rng = np.random.default_rng(17)
    for shape in [(700, 300), (90, 140)]:        # second call smaller, and wide
        a = rng.standard_normal(shape)
        u0, s0, v0, sw0 = _run(lib, a, noalloc=False)
        u1, s1, v1, sw1 = _run(lib, a, noalloc=True)
        assert sw1 == sw0
        np.testing.assert_allclose(s1, s0, rtol=1e-14, atol=0)
        np.testing.assert_allclose(s1, np.linalg.svd(a, compute_uv=False), rtol=1e-12)
        np.testing.assert_allclose(u1, u0, atol=1e-13)
        np.testing.assert_allclose(v1, v0, atol=1e-13)
        np.testing.assert_allclose((u1 * s1) @ v1, a, atol=1e-12 * np.abs(a).max())


def test_noalloc_svd_refuses_sizes_above_the_scratch(lib):
    L = lib
    m, n = 8192, 4096
    need = L.lib().tn_svd_workspace(m, n, L.TN_F64)
    assert need > (64 << 20)
    rows, cols = L.modes([m], [n]), L.modes([n], [1])
    sweeps = ctypes.c_int(0)
    rc = L.lib().tn_svd_noalloc(None, ctypes.byref(rows), ctypes.byref(cols), m, n, None, None,
                                None, L.TN_F64, 0.0, 0, ctypes.byref(sweeps), L.stream_ptr())
    assert rc == L.TN_ERR_WORKSPACE
    assert b"scratch" in L.lib().tn_last_error()
