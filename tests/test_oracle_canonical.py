"""Pins of the canonical-form (Lambda Gamma) MPS oracle (onedim_core.py:880-1326,
1836-1928) by dense linear algebra: Lambdas are the Schmidt values of the
dense state, two-site updates with truncation are the Eckart-Young truncation
of the updated state at that bond, swaps transpose two physical axes, local
expectation values and density matrices equal the dense ones.  The canonical
form from right_canonical_to_canonical drops the sign of the last 1 x 1 SVD
factor (as the reference does), so states are compared through v v^T."""
import numpy as np
import pytest

from oracle import tn_oracle as o
from oracle import models as om


def random_mps(N, d, chi, rng):
    arrs = [rng.standard_normal((d, 1 if i == 0 else chi, 1 if i == N - 1 else chi))
            for i in range(N)]
    return o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrs])


def canonical(N, d, chi, rng, normalise=False):
    m = random_mps(N, d, chi, rng)
    if normalise:
        m[0].data = m[0].data / np.linalg.norm(om.mps_to_dense(m))
    v = om.mps_to_dense(m)
    return o.canonical_form_mps(m), v


def dense_c(c):
    return om.mps_to_dense(o.canonical_to_left_canonical(c))


def proj(v):
    return np.outer(v, v)


def schmidt(v, N, d, b):
    return np.linalg.svd(v.reshape(d ** b, d ** (N - b)), compute_uv=False)


def truncate_at(v, N, d, b, chi):
    u, s, vt = np.linalg.svd(v.reshape(d ** b, d ** (N - b)), full_matrices=False)
    return ((u[:, :chi] * s[:chi]) @ vt[:chi]).reshape(-1)


def gate_on(v, N, d, G, k):
    """(G on sites k, k+1) v; G indexed [out0, out1, in0, in1]."""
    t = v.reshape([d] * N)
    t = np.tensordot(G, t, ([2, 3], [k, k + 1]))
    return np.moveaxis(t, [0, 1], [k, k + 1]).reshape(-1)


def test_canonical_form_lambdas_are_schmidt_values(rng):
    N, d = 7, 2
    c, v = canonical(N, d, 5, rng)
    assert c.nsites == 2 * N + 1 and c.nsites_physical == N
    assert c.check_canonical_form(threshold=1e-10)[:2] == ([], [])  # unnormalised: edges listed
    for b in range(1, N):
        lam = np.diag(c[c.singular_site(b)].data)
        ref = schmidt(v, N, d, b)
        np.testing.assert_allclose(lam, ref[:len(lam)], rtol=1e-10, atol=1e-12 * ref[0])
        assert np.all(ref[len(lam):] < 1e-12 * ref[0])
    np.testing.assert_allclose(c.norm(), np.linalg.norm(v), rtol=1e-12)
    np.testing.assert_allclose(proj(dense_c(c)), proj(v), atol=1e-12 * (v @ v))
    r = o.canonical_to_right_canonical(c)
    np.testing.assert_allclose(proj(om.mps_to_dense(r)), proj(v), atol=1e-12 * (v @ v))
    # the edge Lambdas' norms go into the first tensor, the last Gamma keeps
    # no Lambda: right canonical exactly when the state is normalised
    cn, vn = canonical(N, d, 5, rng, normalise=True)
    assert cn.check_canonical_form(threshold=1e-10) == ([], [], [])
    assert o.canonical_to_right_canonical(cn).check_canonical_form(threshold=1e-10)[1] == 0
    assert o.canonical_to_left_canonical(cn).check_canonical_form(threshold=1e-10)[0] == N - 1
    np.testing.assert_allclose(o.inner_product_mps(c, c), v @ v, rtol=1e-12)
    lc = o.left_canonical_to_canonical(o.left_canonical_form_mps(random_mps(N, d, 3, rng)))
    assert lc.check_canonical_form(threshold=1e-10)[:2] == ([], [])
    with pytest.raises(NotImplementedError):
        o.contract_mps_mpo(c, om.xxz_mpo(N))


@pytest.mark.parametrize("chi", [None, 3])
def test_two_site_gate_is_eckart_young_update(chi, rng):
    N, d, k = 6, 2, 2
    c, v = canonical(N, d, 4, rng)
    G = rng.standard_normal((d, d, d, d))
    gate = o.OTensor(G, ["o0", "o1", "i0", "i1"])
    c.apply_gate(gate, k, chi=chi)
    w = gate_on(v, N, d, G, k)
    ref = w if chi is None else truncate_at(w, N, d, k + 1, chi)
    np.testing.assert_allclose(proj(dense_c(c)), proj(ref), atol=1e-10 * (ref @ ref))
    lam = np.diag(c[c.singular_site(k + 1)].data)
    np.testing.assert_allclose(lam, schmidt(w, N, d, k + 1)[:len(lam)], rtol=1e-10)
    if chi is None:
        # an orthogonal gate keeps the whole chain canonical
        cu, _ = canonical(N, d, 4, rng, normalise=True)
        Q = np.linalg.qr(rng.standard_normal((d * d, d * d)))[0].reshape(d, d, d, d)
        cu.apply_gate(o.OTensor(Q, ["o0", "o1", "i0", "i1"]), k)
        assert cu.check_canonical_form(threshold=1e-10) == ([], [], [])


def test_one_site_gate_and_explicit_labels(rng):
    N, d = 5, 3
    c, v = canonical(N, d, 4, rng)
    G = rng.standard_normal((d, d))
    c.apply_gate(o.OTensor(G, ["x", "y"]), 1, gate_outputs=["x"])
    t = np.moveaxis(np.tensordot(G, v.reshape([d] * N), ([1], [1])), 0, 1).reshape(-1)
    np.testing.assert_allclose(proj(dense_c(c)), proj(t), atol=1e-11 * (t @ t))
    with pytest.raises(NotImplementedError):
        c.apply_gate(o.OTensor(rng.standard_normal((d,) * 6), list("abcdef")), 0)
    with pytest.raises(ValueError):
        c.apply_gate(o.OTensor(G, ["x", "y"]), 0, gate_outputs=["x"], gate_inputs=["x", "y"])


@pytest.mark.parametrize("chi", [None, 2])
def test_swap_gate_transposes_sites(chi, rng):
    N, d, k = 6, 2, 3
    c, v = canonical(N, d, 4, rng)
    c.swap_gate(k, chi=chi)
    t = v.reshape([d] * N)
    w = np.swapaxes(t, k, k + 1).reshape(-1)
    ref = w if chi is None else truncate_at(w, N, d, k + 1, chi)
    np.testing.assert_allclose(proj(dense_c(c)), proj(ref), atol=1e-10 * (ref @ ref))
    assert all(x.labels.count("phys") <= 1 for x in c)


def test_compress_bond_truncates_that_bond(rng):
    N, d = 8, 2
    c, v = canonical(N, d, 6, rng)
    n = 2  # the early-exit test reads position 4n, which exists for n <= N/2
    c.compress_bond(n, chi=2)
    ref = truncate_at(v, N, d, n, 2)
    np.testing.assert_allclose(proj(dense_c(c)), proj(ref), atol=1e-10 * (ref @ ref))
    assert c.singulardim(n) == 2
    c2, _ = canonical(N, d, 6, rng)
    with pytest.raises(IndexError):
        c2.compress_bond(5, chi=2)  # position 20 > 2N: the reference's IndexError
    with pytest.raises(NotImplementedError):
        c2.compress_all()


def test_expval_and_ptrace_dense(rng):
    N, d = 6, 2
    c, v = canonical(N, d, 4, rng)
    Z = np.diag([1.0, -1.0])
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    t = v.reshape([d] * N)
    for k in range(N):
        ev = float(np.squeeze(c.expval(o.OTensor(Z, ["o", "i"]), k).data))
        Zv = np.moveaxis(np.tensordot(Z, t, ([1], [k])), 0, k).reshape(-1)
        np.testing.assert_allclose(ev, v @ Zv, rtol=1e-10, atol=1e-12 * (v @ v))
    G = np.einsum("ac,bd->abcd", X, Z)
    ev2 = float(np.squeeze(c.expval(o.OTensor(G, ["o0", "o1", "i0", "i1"]), 2).data))
    np.testing.assert_allclose(ev2, v @ gate_on(v, N, d, G, 2), rtol=1e-10, atol=1e-12 * (v @ v))
    rho = c.ptrace(1, 2)
    tt = t.reshape(d, d, d, d ** 3)
    ref = np.einsum("aijz,aklz->ijkl", tt, tt)
    outs = [x for x in rho.labels if "_out" in x]
    ins = [x for x in rho.labels if "_in" in x]
    got = np.transpose(rho.data, [rho.labels.index(x) for x in outs + ins])
    np.testing.assert_allclose(got, ref, atol=1e-11 * (v @ v))
