"""Parity of MatrixProductStateCanonical (onedim_core.py:880-1326) and the
canonical-form conversions (onedim_core.py:1448-1454, 1836-1928) with the CPU
oracle, on the GPU; Tensor.inv (tensor.py:487-488) against numpy.linalg.inv.

Bar: singular values (the Lambdas) equal to the oracle's to <= 1e-10
(relative), states compared sign-free through ||v v^T - w w^T|| <= 1e-10
||v||^2 (the canonical form drops the sign of its last 1 x 1 SVD factor, as
the reference does), expectation values and density matrices to <= 1e-10,
labels and bond dimensions identical; diagonal inverses bit-exact."""
import numpy as np
import pytest

from oracle import tn_oracle as o
from oracle import models as om

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def tn():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tncontract_amd
    tncontract_amd.native_library()
    return tncontract_amd


def arrays(N, d, chi, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((d, 1 if i == 0 else chi, 1 if i == N - 1 else chi))
            for i in range(N)]


def both_canonical(tn, N, d, chi, seed):
    arrs = arrays(N, d, chi, seed)
    g = tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"]) for a in arrs])
    r = o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrs])
    return tn.onedim.canonical_form_mps(g), o.canonical_form_mps(r)


def to_oracle(c):
    return o.OMPSC([o.OTensor(t.numpy(), list(t.labels)) for t in c], c.left_label,
                   c.right_label, c.phys_label)


def dense(c):
    return om.mps_to_dense(o.canonical_to_left_canonical(to_oracle(c) if not isinstance(c, o.OMPSC)
                                                         else c))


def same_state(g, r, tol=1e-10):
    a, b = dense(g), dense(r)
    assert np.linalg.norm(np.outer(a, a) - np.outer(b, b)) <= tol * (b @ b)


def same_lambdas(g, r):
    assert g.bonddims() == r.bonddims()
    assert [t.labels for t in g] == [t.labels for t in r]
    for p in range(0, g.nsites, 2):
        lg = np.diag(g[p].numpy())
        lr = np.diag(r[p].numpy() if hasattr(r[p], "numpy") else r[p].data)
        np.testing.assert_allclose(lg, lr, rtol=1e-10, atol=1e-12 * max(abs(lr).max(), 1e-300))


def test_tensor_inv(tn):
    rng = np.random.default_rng(0)
    d = rng.standard_normal(40) + 3.0
    t = tn.Tensor(np.diag(d), ["a", "b"])
    t.inv()
    assert np.array_equal(t.numpy(), np.linalg.inv(np.diag(d)))
    a = rng.standard_normal((30, 30))
    g = tn.Tensor(a, ["a", "b"])
    g.inv()
    np.testing.assert_allclose(g.numpy() @ a, np.eye(30), atol=1e-10)
    np.testing.assert_allclose(g.numpy(), np.linalg.inv(a), rtol=1e-9, atol=1e-11)
    strided = tn.Tensor(a.T.copy(), ["a", "b"])
    strided.move_index("b", 0)  # a view of a^T^T = a
    strided.inv()
    np.testing.assert_allclose(strided.numpy(), np.linalg.inv(a), rtol=1e-9, atol=1e-11)
    with pytest.raises(np.linalg.LinAlgError):
        tn.Tensor(np.diag([1.0, 0.0, 2.0]), ["a", "b"]).inv()
    with pytest.raises(np.linalg.LinAlgError):
        tn.Tensor(np.ones((3, 4)), ["a", "b"]).inv()


@pytest.mark.parametrize("N,d,chi", [(7, 2, 5), (6, 3, 4), (9, 2, 8)])
def test_canonical_form_matches_oracle(tn, N, d, chi):
    g, r = both_canonical(tn, N, d, chi, seed=N + chi)
    same_lambdas(g, r)
    same_state(g, r)
    assert g.nsites_physical == r.nsites_physical == N
    np.testing.assert_allclose(g.norm(), r.norm(), rtol=1e-12)
    assert g.check_canonical_form(threshold=1e-10, print_output=False) == \
        r.check_canonical_form(threshold=1e-10)
    np.testing.assert_allclose(float(tn.onedim.inner_product_mps(g, g)),
                               float(o.inner_product_mps(r, r)), rtol=1e-11)
    for conv, oconv in ((tn.onedim.canonical_to_left_canonical, o.canonical_to_left_canonical),
                        (tn.onedim.canonical_to_right_canonical, o.canonical_to_right_canonical)):
        x, y = conv(g), oconv(r)
        a = om.mps_to_dense(o.OMPS([o.OTensor(t.numpy(), t.labels) for t in x]))
        b = om.mps_to_dense(y)
        assert np.linalg.norm(np.outer(a, a) - np.outer(b, b)) <= 1e-10 * (b @ b)
    lg = tn.onedim.left_canonical_to_canonical(tn.onedim.left_canonical_form_mps(
        tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"])
                                      for a in arrays(N, d, chi, 1)])))
    lr = o.left_canonical_to_canonical(o.left_canonical_form_mps(
        o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrays(N, d, chi, 1)])))
    same_lambdas(lg, lr)
    with pytest.raises(NotImplementedError):
        tn.onedim.contract_mps_mpo(g, tn.onedim.xxz_mpo(N))


@pytest.mark.parametrize("chi", [None, 3])
def test_apply_gate_matches_oracle(tn, chi):
    g, r = both_canonical(tn, 8, 2, 6, seed=3)
    rng = np.random.default_rng(4)
    G2 = rng.standard_normal((2, 2, 2, 2))
    G1 = rng.standard_normal((2, 2))
    for k in (0, 3, 6):
        g.apply_gate(tn.Tensor(G2, ["o0", "o1", "i0", "i1"]), k, chi=chi)
        r.apply_gate(o.OTensor(G2, ["o0", "o1", "i0", "i1"]), k, chi=chi)
        same_lambdas(g, r)
    g.apply_gate(tn.Tensor(G1, ["x", "y"]), 2, gate_outputs=["x"])
    r.apply_gate(o.OTensor(G1, ["x", "y"]), 2, gate_outputs=["x"])
    same_lambdas(g, r)
    same_state(g, r)
    with pytest.raises(NotImplementedError):
        g.apply_gate(tn.Tensor(rng.standard_normal((2,) * 6), list("abcdef")), 0)


@pytest.mark.parametrize("chi", [None, 2])
def test_swap_and_compress_bond_match_oracle(tn, chi):
    g, r = both_canonical(tn, 8, 2, 6, seed=5)
    for k in (0, 4, 6):
        g.swap_gate(k, chi=chi)
        r.swap_gate(k, chi=chi)
        same_lambdas(g, r)
    same_state(g, r)
    g.compress_bond(2, chi=2)
    r.compress_bond(2, chi=2)
    same_lambdas(g, r)
    same_state(g, r)
    with pytest.raises(IndexError):
        g.compress_bond(5, chi=2)
    with pytest.raises(NotImplementedError):
        g.compress_all()


def test_expval_and_ptrace_match_oracle(tn):
    g, r = both_canonical(tn, 7, 2, 4, seed=6)
    Z = np.diag([1.0, -1.0])
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    for k in range(7):
        eg = g.expval(tn.Tensor(Z, ["o", "i"]), k)
        er = r.expval(o.OTensor(Z, ["o", "i"]), k)
        np.testing.assert_allclose(float(eg.numpy().reshape(-1)[0]),
                                   float(np.asarray(er.data).reshape(-1)[0]), rtol=1e-10,
                                   atol=1e-12)
    G = np.einsum("ac,bd->abcd", X, Z)
    eg = g.expval(tn.Tensor(G, ["o0", "o1", "i0", "i1"]), 2)
    er = r.expval(o.OTensor(G, ["o0", "o1", "i0", "i1"]), 2)
    np.testing.assert_allclose(float(eg.numpy().reshape(-1)[0]),
                               float(np.asarray(er.data).reshape(-1)[0]), rtol=1e-10, atol=1e-12)
    rg, rr = g.ptrace(1, 3), r.ptrace(1, 3)
    assert rg.labels == rr.labels
    np.testing.assert_allclose(rg.numpy(), rr.data, rtol=1e-10, atol=1e-12 * np.abs(rr.data).max())


def test_batched_svd_fp64_matches_lapack(tn):
    from tncontract_amd.batched import batched_svd
    rng = np.random.default_rng(9)
    shapes = [(40, 40), (100, 30), (30, 100), (256, 256), (7, 3), (3, 7), (1, 5), (64, 200)]
    mats = [rng.standard_normal(s) for s in shapes]
    mats[3][:, 100:] = mats[3][:, :156] @ rng.standard_normal((156, 156))  # rank-deficient
    U, S, Vt, sweeps = batched_svd([torch.from_numpy(a).cuda() for a in mats])
    for a, u, s, vt in zip(mats, U, S, Vt):
        u, s, vt = u.cpu().numpy(), s.cpu().numpy(), vt.cpu().numpy()
        ref = np.linalg.svd(a, compute_uv=False)
        np.testing.assert_allclose(s, ref, rtol=1e-11, atol=1e-12 * ref[0])
        np.testing.assert_allclose((u * s) @ vt, a, atol=1e-11 * ref[0])
        r = int(np.count_nonzero(ref > 1e-10 * ref[0]))
        np.testing.assert_allclose(u[:, :r].T @ u[:, :r], np.eye(r), atol=1e-11)
        np.testing.assert_allclose(vt[:r] @ vt[:r].T, np.eye(r), atol=1e-11)
    assert np.all(sweeps >= 1)


@pytest.mark.parametrize("chi", [None, 5])
def test_apply_gate_layer_matches_sequential_and_oracle(tn, chi):
    g, r = both_canonical(tn, 10, 2, 8, seed=11)
    seq = g.copy()
    rng = np.random.default_rng(12)
    G = [rng.standard_normal((2, 2, 2, 2)) for _ in range(5)]
    ks = [0, 2, 4, 6, 8]
    g.apply_gate_layer([tn.Tensor(x, ["o0", "o1", "i0", "i1"]) for x in G], ks, chi=chi)
    for x, k in zip(G, ks):
        seq.apply_gate(tn.Tensor(x, ["o0", "o1", "i0", "i1"]), k, chi=chi)
        r.apply_gate(o.OTensor(x, ["o0", "o1", "i0", "i1"]), k, chi=chi)
    same_lambdas(g, r)
    same_lambdas(g, seq)
    same_state(g, r)
    # odd layer with one shared gate, then the canonical form and the oracle again
    one = rng.standard_normal((2, 2, 2, 2))
    g.apply_gate_layer(tn.Tensor(one, ["a", "b", "c", "d"]), [1, 3, 5, 7], chi=chi)
    for k in (1, 3, 5, 7):
        r.apply_gate(o.OTensor(one, ["a", "b", "c", "d"]), k, chi=chi)
    same_lambdas(g, r)
    same_state(g, r)
    with pytest.raises(ValueError):
        g.apply_gate_layer(tn.Tensor(one, ["a", "b", "c", "d"]), [2, 3])


def test_batched_svd_many_single_chunk_jobs(tn):
    """More (job, pair) items than compute units on the one-launch Jacobi
    step (every job within one 256-row chunk): 48 graded 256 x 256 matrices
    and 40 tall 250 x 96 ones, against LAPACK."""
    from tncontract_amd.batched import batched_svd
    rng = np.random.default_rng(21)
    mats = []
    for i in range(48):
        u = np.linalg.qr(rng.standard_normal((256, 256)))[0]
        v = np.linalg.qr(rng.standard_normal((256, 256)))[0]
        mats.append((u * np.logspace(0, -4 - (i % 5), 256)) @ v.T)
    mats += [rng.standard_normal((250, 96)) for _ in range(40)]
    U, S, Vt, sweeps = batched_svd([torch.from_numpy(a).cuda() for a in mats])
    for a, u, s, vt in zip(mats, U, S, Vt):
        u, s, vt = u.cpu().numpy(), s.cpu().numpy(), vt.cpu().numpy()
        ref = np.linalg.svd(a, compute_uv=False)
        np.testing.assert_allclose(s, ref, rtol=1e-11, atol=1e-12 * ref[0])
        np.testing.assert_allclose((u * s) @ vt, a, atol=1e-11 * ref[0])
        np.testing.assert_allclose(vt @ vt.T, np.eye(len(s)), atol=1e-11)


def test_singular_lambda_raises_like_np_linalg_inv(tn):
    """A zero singular value next to the gate's bond: the reference's
    inverse of the outer Lambda (np.linalg.inv, tensor.py:487-488) raises
    LinAlgError, for one gate and for a layer, before the state changes."""
    g, _ = both_canonical(tn, 6, 2, 4, seed=13)
    lam = g[2].numpy().copy()
    lam[-1, -1] = 0.0
    g[2] = tn.Tensor(lam, list(g[2].labels))
    before = [t.numpy().copy() for t in g]
    G = np.eye(4).reshape(2, 2, 2, 2)
    with pytest.raises(np.linalg.LinAlgError):
        g.apply_gate(tn.Tensor(G, ["o0", "o1", "i0", "i1"]), 1)
    with pytest.raises(np.linalg.LinAlgError):
        g.apply_gate_layer(tn.Tensor(G, ["o0", "o1", "i0", "i1"]), [1, 3])
    assert all(np.array_equal(a, t.numpy()) for a, t in zip(before, g))
