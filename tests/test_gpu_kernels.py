"""Parity of the HIP kernels against NumPy FP64 on seeded inputs (GPU)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def L():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from tncontract_amd import _lib
    _lib.lib()
    assert len(_lib.loaded_hip_runtimes()) == 1, _lib.loaded_hip_runtimes()
    return _lib


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def test_permute_matches_numpy(L):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((7, 33, 65, 3))
    for perm in [(0, 1, 2, 3), (3, 2, 1, 0), (0, 2, 1, 3), (2, 0, 3, 1), (1, 3, 0, 2)]:
        t = dev(a).permute(perm)
        out = L.materialize(t)
        np.testing.assert_array_equal(out.cpu().numpy(), np.transpose(a, perm))
    b = rng.standard_normal((130, 70)).astype(np.float32)
    out = L.materialize(dev(b).t())
    np.testing.assert_array_equal(out.cpu().numpy(), b.T)


@pytest.mark.parametrize("shape", [(5, 7, 3), (64, 64, 64), (130, 257, 70), (300, 200, 1000),
                                   (16, 4096, 2), (1024, 1, 2048), (1, 1, 1)])
def test_contract_fp64_matches_numpy(L, shape):
    from tncontract_amd import tensor as T
    M, K, N = shape
    rng = np.random.default_rng(M * 7 + N)
    a = rng.standard_normal((M, K))
    b = rng.standard_normal((K, N))
    c = T._contract_data(dev(a), [1], dev(b), [0]).cpu().numpy()
    ref = a @ b
    assert np.abs(c - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * np.sqrt(K)


def test_contract_permuted_operands_fused(L):
    from tncontract_amd import tensor as T
    rng = np.random.default_rng(5)
    a = rng.standard_normal((6, 40, 5, 30))
    b = rng.standard_normal((30, 7, 6, 9))
    # contract a's axes (3, 0) with b's (0, 2); a and b are also permuted views
    da = dev(np.transpose(a, (2, 0, 3, 1))).permute(1, 3, 0, 2)  # == a as a strided view
    db = dev(np.transpose(b, (1, 3, 0, 2))).permute(2, 0, 3, 1)  # == b
    c = T._contract_data(da, [3, 0], db, [0, 2]).cpu().numpy()
    ref = np.tensordot(a, b, ([3, 0], [0, 2]))
    np.testing.assert_allclose(c, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-5), ("bf16x3", 2e-4)])
def test_contract_mixed_precision(L, prec, tol):
    from tncontract_amd import tensor as T
    rng = np.random.default_rng(11)
    a = rng.standard_normal((300, 513))
    b = rng.standard_normal((513, 250))
    c = T._contract_data(dev(a), [1], dev(b), [0], precision=prec).cpu().numpy().astype(np.float64)
    ref = a @ b
    rel = np.linalg.norm(c - ref) / np.linalg.norm(ref)
    assert rel < tol, rel


@pytest.mark.parametrize("m,n", [(1, 1), (3, 5), (8, 8), (37, 13), (13, 37), (200, 64), (64, 200),
                                 (300, 300), (1280, 640), (2100, 90), (90, 2100)])
def test_qr_matches_lapack(L, m, n):
    from tncontract_amd import tensor as T
    rng = np.random.default_rng(m * 1000 + n)
    a = rng.standard_normal((m, n))
    q, r = T._qr_device(dev(a), 1)
    q, r = q.cpu().numpy(), r.cpu().numpy()
    k = min(m, n)
    assert q.shape == (m, k) and r.shape == (k, n)
    np.testing.assert_allclose(q @ r, a, atol=1e-12 * np.sqrt(n) * np.abs(a).max() * 10)
    np.testing.assert_allclose(q.T @ q, np.eye(k), atol=1e-13 * np.sqrt(m) * 10)
    assert np.all(np.tril(r, -1) == 0)
    # same Householder sign convention as LAPACK: Q and R agree with numpy
    qn, rn = np.linalg.qr(a, mode="reduced")
    np.testing.assert_allclose(r, rn, atol=1e-10 * np.abs(rn).max())


@pytest.mark.parametrize("m,n", [(1, 1), (4, 3), (3, 4), (32, 32), (33, 17), (17, 33), (100, 64),
                                 (64, 100), (256, 640), (640, 256), (300, 300)])
def test_svd_matches_lapack(L, m, n):
    from tncontract_amd import tensor as T
    rng = np.random.default_rng(m * 31 + n)
    a = rng.standard_normal((m, n))
    u, s, vt = T._svd_device(dev(a), 1)
    u, s, vt = u.cpu().numpy(), s.cpu().numpy(), vt.cpu().numpy()
    k = min(m, n)
    sref = np.linalg.svd(a, compute_uv=False)
    np.testing.assert_allclose(s, sref, rtol=1e-12, atol=1e-13 * sref[0])
    np.testing.assert_allclose((u * s) @ vt, a, atol=1e-12 * sref[0] * np.sqrt(k))
    np.testing.assert_allclose(u.T @ u, np.eye(k), atol=1e-12 * np.sqrt(k))
    np.testing.assert_allclose(vt @ vt.T, np.eye(k), atol=1e-12 * np.sqrt(k))


def test_svd_graded_spectrum(L):
    from tncontract_amd import tensor as T
    rng = np.random.default_rng(3)
    m, n = 200, 120
    uu, _ = np.linalg.qr(rng.standard_normal((m, n)))
    vv, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.exp(-np.arange(n) / 8.0)
    a = (uu * s) @ vv.T
    u, sg, vt = T._svd_device(dev(a), 1)
    sg = sg.cpu().numpy()
    np.testing.assert_allclose(sg, s, rtol=1e-9, atol=1e-14)


def test_eigh_psd_ties_across_sort_tiles(L):
    """The device rank sort with repeated keys and more keys than one LDS
    tile (2304 > 2048): the PSD eigensolver of a diagonal matrix whose 2304
    entries take 24 distinct values sorts them twice (the static pivot and
    the final order).  Eigenvalues must be the sorted diagonal and U a
    signed permutation matching them."""
    from tncontract_amd import tensor as T
    n = 2304
    rng = np.random.default_rng(5)
    diag = rng.permutation(np.repeat(np.arange(1.0, 25.0), n // 24))
    u, lam = T._eigh_psd_device(dev(np.diag(diag)), 1)
    u, lam = u.cpu().numpy(), lam.cpu().numpy()
    np.testing.assert_allclose(lam, np.sort(diag)[::-1], rtol=1e-14, atol=0)
    au = np.abs(u)
    np.testing.assert_allclose(au.max(axis=0), np.ones(n), rtol=0, atol=1e-14)
    np.testing.assert_allclose(au.sum(axis=0), np.ones(n), rtol=0, atol=1e-12)
    np.testing.assert_array_equal(diag[np.argmax(au, axis=0)], np.sort(diag)[::-1])
    assert len(set(np.argmax(au, axis=0).tolist())) == n  # a permutation


def test_svd_repeated_singular_values(L):
    """Equal singular values (a scaled orthogonal block beside a graded one)
    through the Jacobi's per-sweep norm sort and the final rank sort."""
    from tncontract_amd import tensor as T
    rng = np.random.default_rng(6)
    m, n = 700, 520
    uu, _ = np.linalg.qr(rng.standard_normal((m, n)))
    vv, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.concatenate([np.full(300, 2.0), np.exp(-np.arange(220) / 40.0)])
    a = (uu * s) @ vv.T
    u, sg, vt = T._svd_device(dev(a), 1)
    u, sg, vt = u.cpu().numpy(), sg.cpu().numpy(), vt.cpu().numpy()
    np.testing.assert_allclose(sg, np.sort(s)[::-1], rtol=1e-12, atol=1e-13 * 2.0)
    np.testing.assert_allclose((u * sg) @ vt, a, atol=1e-12 * 2.0 * np.sqrt(n))
    np.testing.assert_allclose(u.T @ u, np.eye(n), atol=1e-12 * np.sqrt(n))


def test_sumsq_and_scale(L):
    rng = np.random.default_rng(9)
    a = rng.standard_normal(1000003)
    t = dev(a)
    assert abs(L.sumsq(t) - np.sum(a * a)) < 1e-9 * np.sum(a * a)
    L.scale_(t, 0.5)
    np.testing.assert_allclose(t.cpu().numpy(), 0.5 * a)


@pytest.mark.parametrize("prec,tol", [("native", 1e-12), ("fp32", 2e-5)])
def test_contract_with_non_fusing_operands(L, prec, tol):
    # environment-style products whose contracted axes are not neighbours in
    # one operand (GEMM-sized, so the operand is first copied into dense
    # (free, contracted) order): F = E(a,al,a',al') . C(a,k,al,t) over (a, al),
    # then K = C . F over (a, al), with C on the left
    from tncontract_amd import tensor as T
    rng = np.random.default_rng(7)
    E = rng.standard_normal((128, 5, 128, 5))
    C = rng.standard_normal((128, 128, 5, 2))
    F_ref = np.tensordot(E, C, axes=([0, 1], [0, 2]))
    F = T._contract_data(dev(E), [0, 1], dev(C), [0, 2], prec)
    scale = np.abs(F_ref).max()
    assert F.shape == F_ref.shape
    assert np.abs(F.double().cpu().numpy() - F_ref).max() <= tol * scale * 10
    K_ref = np.tensordot(C, F_ref, axes=([0, 2], [0, 1]))
    K = T._contract_data(dev(C), [0, 2], dev(F_ref), [0, 1], prec)
    assert K.shape == K_ref.shape
    assert np.abs(K.double().cpu().numpy() - K_ref).max() <= tol * np.abs(K_ref).max() * 10
    # a strided (transposed) view as the operand
    Ct = dev(C.transpose(1, 0, 3, 2).copy()).permute(1, 0, 3, 2)
    F2 = T._contract_data(dev(E), [0, 1], Ct, [0, 2], prec)
    assert np.abs(F2.double().cpu().numpy() - F_ref).max() <= tol * scale * 10
