"""Cholesky QR with a Newton-Schulz correction (tn_cholqr), the
factorisation variational_compress uses for its site updates in place of
tensor_lq (tensor.py:1058-1109, onedim_core.py:560-631), against
numpy.linalg.qr on the same matrices.

Bar: Q R = A to 1e-13 relative, max |Q^T Q - I| <= 1e-12, the column space
of Q equal to LAPACK's (|Q Q^T - Ql Ql^T| <= 1e-12), R equal to LAPACK's R
up to the sign of each row to max(1e-11, 1e-15 cond^2) of its largest
entry (R = Q^T A is triangular up to O(cond^2 eps)); inputs too
ill-conditioned for Cholesky QR are refused and the caller's fallback
(tensor_lq, Householder) gives tensor_lq's result exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def tn():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tncontract_amd
    tncontract_amd.native_library()
    return tncontract_amd


def graded(rng, m, n, cond):
    u = np.linalg.qr(rng.standard_normal((m, n)))[0]
    v = np.linalg.qr(rng.standard_normal((n, n)))[0]
    return (u * np.logspace(0, -np.log10(cond), n)) @ v.T


@pytest.mark.parametrize("m,n,cond", [(1024, 512, 1e2), (300, 200, 1e4), (64, 64, 10.0),
                                      (1000, 70, 1e3), (129, 1, 1.0), (513, 65, 1e4)])
def test_cholqr_matches_lapack_up_to_row_signs(tn, m, n, cond):
    from tncontract_amd.tensor import _cholqr_device
    rng = np.random.default_rng(m + n)
    a = graded(rng, m, n, cond)
    res = _cholqr_device(torch.from_numpy(a).cuda(), 1)
    assert res is not None
    q, r = (x.cpu().numpy() for x in res)
    assert q.shape == (m, n) and r.shape == (n, n)
    np.testing.assert_allclose(q @ r, a, atol=1e-13 * np.abs(a).max() * np.sqrt(n))
    assert np.abs(q.T @ q - np.eye(n)).max() <= 1e-12
    ql, rl = np.linalg.qr(a)
    assert np.abs(q @ q.T - ql @ ql.T).max() <= 1e-12
    assert np.all(np.diag(r) > 0)
    tol = max(1e-11, 1e-15 * cond ** 2)  # R is triangular up to O(cond^2 eps)
    np.testing.assert_allclose(np.abs(r), np.abs(rl), rtol=0, atol=tol * np.abs(rl).max())


def test_cholqr_strided_view(tn):
    from tncontract_amd.tensor import _cholqr_device
    rng = np.random.default_rng(5)
    a = rng.standard_normal((40, 6, 30))  # rows (a, c), cols b after a permuted view
    d = torch.from_numpy(a).cuda().permute(0, 2, 1)
    q, r = (x.cpu().numpy() for x in _cholqr_device(d, 2))
    mat = a.transpose(0, 2, 1).reshape(1200, 6)
    np.testing.assert_allclose(q @ r, mat, atol=1e-13 * np.abs(mat).max() * 4)
    assert np.abs(q.T @ q - np.eye(6)).max() <= 1e-12


def test_cholqr_refuses_ill_conditioned_and_lq_falls_back(tn):
    from tncontract_amd.tensor import _cholqr_device, _lq_rowspace, tensor_lq
    rng = np.random.default_rng(7)
    a = graded(rng, 400, 100, 1e12)
    assert _cholqr_device(torch.from_numpy(a).cuda(), 1) is None
    sing = a.copy()
    sing[:, 50] = sing[:, 3]  # exactly rank-deficient
    assert _cholqr_device(torch.from_numpy(sing).cuda(), 1) is None
    # the variational LQ of the ill-conditioned block is tensor_lq's, exactly
    t = tn.Tensor(a.T.reshape(100, 20, 20), ["l", "p", "r"])
    L1, Q1 = _lq_rowspace(t, "l", lq_label="z")
    L2, Q2 = tensor_lq(t, "l", lq_label="z")
    assert L1.labels == L2.labels and Q1.labels == Q2.labels
    assert np.array_equal(L1.numpy(), L2.numpy()) and np.array_equal(Q1.numpy(), Q2.numpy())
    # a wide block (rows > product of the other dims) also falls back
    w = tn.Tensor(rng.standard_normal((30, 2, 4)), ["l", "p", "r"])
    L3, Q3 = _lq_rowspace(w, "l", lq_label="z")
    L4, Q4 = tensor_lq(w, "l", lq_label="z")
    assert np.array_equal(L3.numpy(), L4.numpy()) and np.array_equal(Q3.numpy(), Q4.numpy())


def test_lq_rowspace_labels_and_product(tn):
    from tncontract_amd.tensor import _lq_rowspace, tensor_lq
    rng = np.random.default_rng(8)
    arr = rng.standard_normal((12, 2, 16))
    t = tn.Tensor(arr, ["left", "phys", "right"])
    Lr, Qr = _lq_rowspace(t, "left", lq_label="z")
    Lh, Qh = tensor_lq(t, "left", lq_label="z")
    assert Lr.labels == Lh.labels and Qr.labels == Qh.labels and Lr.shape == Lh.shape
    lr, qr_ = Lr.numpy(), Qr.numpy()
    lh, qh = Lh.numpy(), Qh.numpy()
    # L Q reproduces the tensor; rows of Q agree with Householder's up to sign
    prod = np.einsum("zl,prz->lpr", lr, qr_)
    np.testing.assert_allclose(prod, arr, atol=1e-13 * np.abs(arr).max())
    sg = np.sign(np.diag(lr.reshape(lr.shape[0], -1)[:, :12]) * np.diag(lh.reshape(lh.shape[0], -1)[:, :12]))
    np.testing.assert_allclose(qr_ * sg, qh, atol=1e-10)
