"""Pins for the CPU oracle (oracle/tn_oracle.py, oracle/models.py).

Each check is chosen so that a plausible mistake in the oracle (a wrong axis
order, a dropped truncation term, a transposed operand) fails it: known
answers from the reference README/docstrings, brute force with np.einsum and
dense state vectors, closed forms, and the invariants of the reference's own
test file on the reference's own fixture.
"""
import numpy as np
import pytest

from oracle import tn_oracle as o
from oracle import models as om

from cases import decaying_sigma, schmidt_chain, schmidt_pair


def T(shape, labels, rng):
    return o.OTensor(rng.standard_normal(shape), labels)


# ---------------------------------------------------------------- axis ops
def test_move_indices_docstring_known_answer(rng):
    # tensor.py:408-441
    t = T((2, 3, 4, 5, 6), ["a", "b", "c", "b", "d"], rng)
    t.move_indices(["d", "b", "c"], 0, preserve_relative_order=True)
    assert t.labels == ["b", "c", "b", "d", "a"] and t.shape == (3, 4, 5, 6, 2)
    t = T((2, 3, 4, 5, 6), ["a", "b", "c", "b", "d"], rng)
    src = t.data.copy()
    t.move_indices(["d", "b", "c"], 0, preserve_relative_order=False)
    assert t.labels == ["d", "b", "b", "c", "a"] and t.shape == (6, 3, 5, 4, 2)
    assert np.array_equal(t.data, np.transpose(src, (4, 1, 3, 2, 0)))


def test_consolidate_indices_fuses_sorted(rng):
    # tensor.py:340-370: alphabetical, duplicates fused with first occurrence slowest
    t = T((2, 3, 4, 5), ["z", "a", "z", "b"], rng)
    src = t.data.copy()
    t.consolidate_indices()
    assert t.labels == ["a", "b", "z"] and t.shape == (3, 5, 8)
    ref = np.transpose(src, (1, 3, 0, 2)).reshape(3, 5, 8)
    assert np.array_equal(t.data, ref)


def test_move_index_is_moveaxis(rng):
    t = T((2, 3, 4, 5), ["a", "b", "c", "d"], rng)
    src = t.data.copy()
    t.move_index("a", 2)
    assert t.labels == ["b", "c", "a", "d"]
    assert np.array_equal(t.data, np.moveaxis(src, 0, 2))
    t.move_index("d", 0)
    assert t.labels == ["d", "b", "c", "a"]


# ---------------------------------------------------------------- contract
def test_contract_docstring_known_answer(rng):
    # tensor.py:662-683
    A = T((2, 2), ["spam", "eggs"], rng)
    B = T((2, 3, 2, 4), [], rng)
    assert B.labels == ["i0", "i1", "i2", "i3"]
    C = o.contract(A, B, "spam", "i2")
    assert C.shape == (2, 2, 3, 4) and C.labels == ["eggs", "i0", "i1", "i3"]
    np.testing.assert_allclose(C.data, np.einsum("se,absd->eabd", A.data, B.data), rtol=1e-13)
    D = o.contract(A, B, ["spam", "eggs"], ["i0", "i2"])
    assert D.shape == (3, 4) and D.labels == ["i1", "i3"]
    np.testing.assert_allclose(D.data, np.einsum("se,sbed->bd", A.data, B.data), rtol=1e-13)
    D2 = A["spam", "eggs"] * B["i0", "i2"]
    np.testing.assert_array_equal(D2.data, D.data)


def test_contract_duplicate_labels_and_slice(rng):
    # all axes with a listed label are contracted (tensor.py:702-712); index_slice selects
    A = T((3, 4, 3), ["x", "y", "x"], rng)
    B = T((3, 5, 3), ["x", "z", "x"], rng)
    C = o.contract(A, B, "x", "x")
    np.testing.assert_allclose(C.data, np.einsum("ayb,azb->yz", A.data, B.data), rtol=1e-13)
    C2 = o.contract(A, B, "x", "x", index_slice1=[-1], index_slice2=[0])
    assert C2.labels == ["x", "y", "z", "x"]
    np.testing.assert_allclose(C2.data, np.einsum("ayk,kzb->ayzb", A.data, B.data), rtol=1e-13)


def test_contract_dimension_mismatch_raises(rng):
    A = T((2, 3), ["a", "b"], rng)
    B = T((4, 3), ["c", "d"], rng)
    with pytest.raises(ValueError):
        o.contract(A, B, "a", "c")


def test_tensor_product_and_trace(rng):
    A = T((2, 3), ["a", "b"], rng)
    B = T((4,), ["c"], rng)
    P = o.tensor_product(A, B)
    assert P.labels == ["a", "b", "c"]
    np.testing.assert_allclose(P.data, np.einsum("ab,c->abc", A.data, B.data))
    C = T((3, 4, 3), ["p", "q", "r"], rng)
    C.trace("p", "r")
    assert C.labels == ["q"]


# ---------------------------------------------------------------- con
def test_con_readme_examples(rng):
    # tncon.py:26-65 / README.md:81-127
    A = T((3, 2, 4), ["a", "b", "c"], rng)
    B = T((3, 4), ["d", "e"], rng)
    C = T((5, 5, 2), ["f", "g", "h"], rng)
    r = o.con(A, B, ("a", "d"), ("c", "e"))
    assert r.labels == ["b"]
    np.testing.assert_allclose(r.data, np.einsum("abc,ac->b", A.data, B.data), rtol=1e-13)
    r = o.con(C, ("f", "g"))
    assert r.labels == ["h"]
    np.testing.assert_allclose(r.data, np.einsum("ffh->h", C.data), rtol=1e-13)
    r = o.con(A, B)
    assert r.shape == (3, 2, 4, 3, 4) and r.labels == ["a", "b", "c", "d", "e"]
    r = o.con(A, B, C, ("a", "d"), ("c", "e"), ("f", "g"), ("h", "b"))
    ref = np.einsum("abc,ac,ffb->", A.data, B.data, C.data)
    np.testing.assert_allclose(r.data, ref, rtol=1e-12)


def test_con_ring_closed_form(rng):
    N = 40
    a = rng.standard_normal((2, 2)) / 1.5
    A = o.OTensor(a, ["left", "right"])
    ts = [A.suf(str(i)) for i in range(N)]
    pairs = [("right" + str(j), "left" + str(j + 1)) for j in range(N - 1)]
    r = o.con(ts, pairs, ("right" + str(N - 1), "left0"))
    np.testing.assert_allclose(r.data, np.trace(np.linalg.matrix_power(a, N)), rtol=1e-10)


def test_con_rejects_duplicates(rng):
    A = T((2, 2), ["a", "b"], rng)
    B = T((2, 2), ["a", "c"], rng)
    with pytest.raises(ValueError):
        o.con(A, B, ("b", "c"))


# ---------------------------------------------------------------- decompositions
def test_tensor_svd_docstring_and_reconstruction(rng):
    a = T((2, 3, 4), ["i0", "i1", "i2"], rng)
    U, S, V = o.tensor_svd(a, ["i0", "i2"])
    assert U.shape == (2, 4, 3) and U.labels == ["i0", "i2", "svd_in"]
    assert V.shape == (3, 3) and V.labels == ["svd_out", "i1"]
    assert S.shape == (3, 3) and S.labels == ["svd_out", "svd_in"]
    b = o.contract(U, o.contract(S, V, "svd_in", "svd_out"), "svd_in", "svd_out")
    assert o.distance(a, b) < 1e-13
    mat = np.transpose(a.data, (0, 2, 1)).reshape(8, 3)
    np.testing.assert_allclose(np.diag(S.data), np.linalg.svd(mat, compute_uv=False), rtol=1e-13)


def test_tensor_qr_docstring_and_reconstruction(rng):
    t = T((2, 3, 4), [], rng)
    Q, R = o.tensor_qr(t, ["i0", "i2"])
    assert Q.shape == (2, 4, 3) and Q.labels == ["i0", "i2", "qr_in"]
    assert R.shape == (3, 3) and R.labels == ["qr_out", "i1"]
    x = o.contract(Q, R, "qr_in", "qr_out")
    assert x.labels == ["i0", "i2", "i1"] and o.distance(x, t) < 1e-13
    q = Q.data.reshape(8, 3)
    np.testing.assert_allclose(q.T @ q, np.eye(3), atol=1e-14)
    assert np.allclose(np.tril(R.data, -1), 0)


def test_truncated_svd_selection_rule(rng):
    # closed form: cap at chi, then threshold; truncated_evals = [thr-cut, chi-cut]
    u, _ = np.linalg.qr(rng.standard_normal((12, 6)))
    v, _ = np.linalg.qr(rng.standard_normal((10, 6)))
    s = np.array([5.0, 2.0, 1.0, 1e-3, 1e-9, 1e-12])
    t = o.OTensor((u * s) @ v.T, ["r", "c"])
    U, V, ev = o.truncated_svd(t, ["r"], chi=5, threshold=1e-6)
    assert U.shape == (12, 4) and V.shape == (4, 10)
    assert len(ev) == 6  # 1 thresholded out of the first 5, then sv[5:] (5 values)
    np.testing.assert_allclose(ev[:2], [1e-9, 1e-12], rtol=1e-3)
    assert np.all(ev[2:] < 1e-14)
    U, V, ev = o.truncated_svd(t, ["r"], chi=0, threshold=1e-4, absolute=False)
    assert U.shape[1] == 4  # sigma/sigma0 > 1e-4 keeps 5,2,1,1e-3
    U, S, V = o.truncated_svd(t, ["r"], chi=2, absorb_singular_values=None)
    np.testing.assert_allclose(np.diag(S.data), [5.0, 2.0], rtol=1e-12)


# ---------------------------------------------------------------- one-dimensional
def random_omps(N, d, chi, rng):
    arrs = [rng.standard_normal((d, 1 if i == 0 else chi, 1 if i == N - 1 else chi)) for i in range(N)]
    return o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrs])


@pytest.mark.parametrize("chi", [2, 4, 8])
def test_svd_compress_equals_tt_rounding_of_dense_state(chi, rng):
    N, d = 9, 2
    m = random_omps(N, d, 6, rng)
    v = om.mps_to_dense(m)
    m.svd_compress(chi=chi, threshold=1e-15)
    ref = om.tt_truncate_dense(v, N, d, chi)
    got = om.mps_to_dense(m)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-11
    # result is right canonical: all sites but site 0 are right isometries
    l, r = m.check_canonical_form(threshold=1e-10)
    assert r == 0
    assert max(m.bonddims()) <= chi


def test_svd_compress_output_axis_order(rng):
    m = random_omps(5, 2, 3, rng)
    m.svd_compress(chi=2)
    assert m[0].labels == ["right", "phys", "left"]
    assert all(t.labels == ["phys", "right", "left"] for t in m.data[1:])


def test_left_canonise_isometry_and_norm(rng):
    m = random_omps(8, 3, 5, rng)
    n0 = np.sqrt(o.inner_product_mps(m, m))
    a = m.copy(); a.left_canonise(qr_decomposition=True)
    assert a.check_canonical_form(threshold=1e-10)[0] == len(a) - 1
    np.testing.assert_allclose(a.norm("left"), n0, rtol=1e-12)
    b = m.copy(); b.left_canonise()
    np.testing.assert_allclose(b.norm("left"), n0, rtol=1e-12)
    c = m.copy(); c.left_canonise(normalise=True)
    np.testing.assert_allclose(c.norm("left"), 1.0, rtol=1e-12)
    np.testing.assert_allclose(om.mps_to_dense(a), om.mps_to_dense(m), atol=1e-12 * n0)


def test_zero_state_special_case(rng):
    m = random_omps(4, 2, 3, rng)
    for t in m:
        t.data[:] = 0
    m.left_canonise()
    assert all(t.shape == (2, 1, 1) for t in m)
    assert all(np.all(t.data == 0) for t in m)


def test_contract_mps_mpo_matches_dense(rng):
    N = 7
    mpo = om.xxz_mpo(N, Jz=1.0, Jxy=1.0, hz=0.5)
    H = om.xxz_dense_hamiltonian(N, 1.0, 1.0, 0.5)
    np.testing.assert_allclose(om.mpo_to_dense(mpo), H, atol=1e-14)
    m = random_omps(N, 2, 4, rng)
    out = o.contract_mps_mpo(m, mpo)
    assert out.phys_label == "physout"
    assert out[1].labels == ["left", "physout", "right"] and out[1].shape == (20, 2, 20)
    np.testing.assert_allclose(om.mps_to_dense(out), H @ om.mps_to_dense(m), atol=1e-12)


def test_xxz_polarised_energy_closed_form():
    N, Jz, Jxy, hz = 8, 1.0, 1.0, 0.5
    up = np.zeros(2 ** N); up[0] = 1.0
    H = om.xxz_dense_hamiltonian(N, Jz, Jxy, hz)
    assert abs(up @ H @ up - ((N - 1) * Jz / 4 + N * hz / 2)) < 1e-13


def test_inner_product_and_distance_dense(rng):
    a = random_omps(6, 2, 3, rng)
    b = random_omps(6, 2, 3, rng)
    va, vb = om.mps_to_dense(a), om.mps_to_dense(b)
    np.testing.assert_allclose(o.inner_product_mps(a, b), va @ vb, rtol=1e-12)
    np.testing.assert_allclose(o.frob_distance_squared(a, b), np.sum((va - vb) ** 2), rtol=1e-11)


def test_expvals_dense(rng):
    N = 6
    m = random_omps(N, 2, 3, rng)
    v = om.mps_to_dense(m)
    v = v / np.linalg.norm(v)
    m2 = o.OMPS([t * (1.0 / np.linalg.norm(om.mps_to_dense(m)) ** (1.0 / N)) for t in m])
    op = o.OTensor(om.SZ, ["out", "in"])
    ev = o.expvals_mps(m2, op)
    ref = [v @ om._site_op(om.SZ, i, N) @ v for i in range(N)]
    np.testing.assert_allclose(ev.real, ref, atol=1e-12)


# ---------------------------------------------------------------- reference fixture
def fixture_omps(arrays, meta):
    return o.OMPS([o.OTensor(a, lab) for a, lab in zip(arrays, meta["labels"])],
                  meta["left_label"], meta["right_label"], meta["phys_label"])


def test_reference_fixture_invariants(fixture_mps_arrays):
    # tncontract/tests/test_canonical_form_mps_conversion.py:31-64 invariants
    arrays, meta = fixture_mps_arrays
    psi = fixture_omps(arrays, meta)
    n2 = o.inner_product_mps(psi, psi)
    np.testing.assert_allclose(n2, 0.4668553069796, atol=1e-12)
    psi.svd_compress(threshold=1e-12, normalise=False)
    np.testing.assert_almost_equal(psi.norm(), np.sqrt(n2), decimal=10)
    np.testing.assert_almost_equal(psi.norm("right"), np.sqrt(n2), decimal=10)
    assert psi.check_canonical_form(threshold=1e-10)[1] == 0
    psi.left_canonise()
    assert psi.check_canonical_form(threshold=1e-10)[0] == len(psi) - 1
    np.testing.assert_almost_equal(psi.norm(), np.sqrt(n2), decimal=10)
    psi.svd_compress(threshold=1e-12, normalise=True)
    np.testing.assert_almost_equal(psi.norm(), 1.0, decimal=10)
    psi2 = fixture_omps(arrays, meta)
    np.testing.assert_almost_equal(o.inner_product_mps(psi, psi2) ** 2 / n2, 1.0, decimal=10)


# ---------------------------------------------------------------- LQ, ladders, variational
def test_tensor_lq_reconstruction_and_triangular(rng):
    t = T((3, 4, 5), ["a", "b", "c"], rng)
    Lt, Q = o.tensor_lq(t, ["a", "b"], lq_label="z")
    # L is the R factor of a QR of the (c | a b) matrix, Q its Q factor
    assert Lt.labels == ["zin", "a", "b"] and Q.labels == ["c", "zout"]
    l = Lt.data.reshape(-1, 12).T
    q = Q.data.T
    # rows of q orthonormal, l lower triangular, l q == the (a b | c) matrix
    np.testing.assert_allclose(q @ q.T, np.eye(q.shape[0]), atol=1e-13)
    assert np.allclose(np.triu(l[: l.shape[1]], 1), 0.0)
    np.testing.assert_allclose(l @ q, t.data.reshape(12, 5), atol=1e-13)
    Lt2, Q2 = o.tensor_lq(t, ["c"], lq_label="z")
    np.testing.assert_allclose(_labelled(o.contract(Lt2, Q2, "zin", "zout"), ["a", "b", "c"]),
                               t.data, atol=1e-13)


def _transfer_chain(a, b, lo, hi):
    """Dense product of the bra/ket transfer matrices of sites lo..hi:
    E[l1, l2, r1, r2] (site arrays ordered (phys, left, right))."""
    E = None
    for i in range(lo, hi + 1):
        A = om._site_array(a[i], a.phys_label, a.left_label, a.right_label)
        B = om._site_array(b[i], b.phys_label, b.left_label, b.right_label)
        e = np.einsum("apb,cpd->acbd", A, B)
        E = e if E is None else np.einsum("acbd,bdef->acef", E, e)
    return E


def _labelled(t, labels):
    return np.transpose(t.data, [t.labels.index(x) for x in labels])


def test_ladder_contract_branches_match_transfer_products(rng):
    N = 6
    a = random_omps(N, 2, 3, rng)
    b = random_omps(N, 2, 4, rng)
    full = _transfer_chain(a, b, 0, N - 1)[0, 0, 0, 0]
    np.testing.assert_allclose(float(o.ladder_contract(a, b, "phys", "phys").data), full, rtol=1e-12)
    # left branch, partial end, intermediates
    inter = o.ladder_contract(a, b, "phys", "phys", end=3, return_intermediate_contractions=True)
    assert len(inter) == 4
    for k, t in enumerate(inter):
        np.testing.assert_allclose(_labelled(t, ["right1", "right2"]),
                                   _transfer_chain(a, b, 0, k)[0, 0], rtol=1e-12, atol=1e-12)
    allint = o.ladder_contract(a, b, "phys", "phys", return_intermediate_contractions=True)
    np.testing.assert_allclose(allint[-1].data.item(), full, rtol=1e-12)
    # right branch (interval ends at the last site)
    r = o.ladder_contract(a, b, "phys", "phys", start=2, end=N - 1)
    np.testing.assert_allclose(_labelled(r, ["left1", "left2"]),
                               _transfer_chain(a, b, 2, N - 1)[:, :, 0, 0], rtol=1e-12, atol=1e-12)
    rint = o.ladder_contract(a, b, "phys", "phys", start=2, end=N - 1,
                             return_intermediate_contractions=True)
    for k, t in zip(range(2, N), rint):
        np.testing.assert_allclose(_labelled(t, ["left1", "left2"]),
                                   _transfer_chain(a, b, k, N - 1)[:, :, 0, 0], rtol=1e-12, atol=1e-12)
    # middle branch
    m = o.ladder_contract(a, b, "phys", "phys", start=1, end=3)
    np.testing.assert_allclose(_labelled(m, ["left1", "left2", "right1", "right2"]),
                               _transfer_chain(a, b, 1, 3), rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError):
        o.ladder_contract(a, b, "phys", "phys", start=3, end=2)


def _xxz_target(N, chi, rng):
    m = random_omps(N, 2, chi, rng)
    return o.contract_mps_mpo(m, om.xxz_mpo(N, 1.0, 0.7, 0.3))


def test_variational_compress_exact_when_chi_suffices(rng):
    N = 6
    tgt = _xxz_target(N, 3, rng)
    v = om.mps_to_dense(tgt)
    out = tgt.variational_compress(8, max_iter=20, tolerance=1e-8)  # 2^3 = exact bond bound
    assert out.phys_label == tgt.phys_label and out.left_label == tgt.left_label
    np.testing.assert_allclose(om.mps_to_dense(out), v, atol=1e-10 * np.linalg.norm(v))
    assert out.check_canonical_form(threshold=1e-10)[0] == N - 1


@pytest.mark.parametrize("chi", [2, 4])
def test_variational_compress_projection_identities(chi, rng):
    N = 9
    tgt = _xxz_target(N, 4, rng)
    v = om.mps_to_dense(tgt)
    s = tgt.copy(); s.svd_compress(chi=chi)
    out = o.variational_compress_mps(tgt, chi, max_iter=60, tolerance=1e-8)
    w = om.mps_to_dense(out)
    d_var = np.linalg.norm(v - w)
    # never worse than its svd_compress starting point; bonds stay within chi
    assert d_var <= np.linalg.norm(v - om.mps_to_dense(s)) * (1 + 1e-12)
    assert max(out.bonddims()) <= chi
    # left canonical; the last site is the orthogonal projection of the target
    # onto the span fixed by the other sites: site == (left isometry)^T target,
    # hence <w, v> == |w|^2 and |v - w|^2 == |v|^2 - |w|^2
    assert out.check_canonical_form(threshold=1e-10)[0] == N - 1
    np.testing.assert_allclose(w @ v, w @ w, rtol=1e-9)
    np.testing.assert_allclose(d_var ** 2, v @ v - w @ w, rtol=1e-7, atol=1e-12 * (v @ v))
    # the functional variant and the method agree
    out2 = tgt.variational_compress(chi, max_iter=60, tolerance=1e-8)
    np.testing.assert_allclose(om.mps_to_dense(out2), w, atol=1e-12 * np.linalg.norm(w))


def test_variational_compress_target_unchanged_and_nonconvergence(rng):
    tgt = _xxz_target(7, 3, rng)
    v = om.mps_to_dense(tgt)
    tgt.variational_compress(3, max_iter=10, tolerance=1e-6)
    np.testing.assert_allclose(om.mps_to_dense(tgt), v, atol=1e-12 * np.linalg.norm(v))
    with pytest.raises(RuntimeError):
        tgt.variational_compress(2, max_iter=1, tolerance=1e-30)


# ---------------------------------------------------------------- tensor <-> matrix
def test_tensor_to_matrix_known_answer(rng):
    # tensor.py:805-818: the row labels are moved to the front in the given
    # order (one move_index each), the rest keep their relative order; the
    # row index runs fastest over the LAST row label.  Checked element by
    # element with explicit index loops (no reshape).
    t = T((2, 3, 4, 5), ["a", "b", "c", "d"], rng)
    M = o.tensor_to_matrix(t, ["c", "a"])
    assert M.shape == (8, 15)
    for a in range(2):
        for b in range(3):
            for c in range(4):
                for d in range(5):
                    assert M[c * 2 + a, b * 5 + d] == t.data[a, b, c, d]
    # the input tensor is not modified
    assert t.labels == ["a", "b", "c", "d"] and t.shape == (2, 3, 4, 5)
    # no row labels: one row; every label a row label: one column
    assert o.tensor_to_matrix(t, []).shape == (1, 120)
    M2 = o.tensor_to_matrix(t, ["d", "c", "b", "a"])
    assert M2.shape == (120, 1)
    assert M2[((4 * 4 + 3) * 3 + 2) * 2 + 1, 0] == t.data[1, 2, 3, 4]


def test_matrix_to_tensor_known_answer(rng):
    # tensor.py:821-827: row-major reshape, default labels i0, i1, ...
    m = rng.standard_normal((6, 4))
    t = o.matrix_to_tensor(m, (2, 3, 4))
    assert t.labels == ["i0", "i1", "i2"] and t.shape == (2, 3, 4)
    for i in range(2):
        for j in range(3):
            for k in range(4):
                f = (i * 3 + j) * 4 + k           # row-major flat position
                assert t.data[i, j, k] == m[f // 4, f % 4]
    t2 = o.matrix_to_tensor(m, (3, 8), ["x", "y"])
    assert t2.labels == ["x", "y"] and t2.data[2, 7] == m[5, 3]
    # round trip with tensor_to_matrix
    back = o.tensor_to_matrix(o.matrix_to_tensor(m, (2, 3, 4), ["p", "q", "r"]), ["p", "q"])
    assert np.array_equal(back, m)


# ---------------------------------------------------------------- keep rule of left_canonise
@pytest.mark.parametrize("scale,threshold,rel_keep,abs_keep", [
    (1e-3, 1e-10, 3, 2),      # relative keeps more than an absolute rule would
    (1e3, 1e-14, 3, 4),       # relative keeps fewer than an absolute rule would
])
def test_left_canonise_threshold_is_relative(scale, threshold, rel_keep, abs_keep):
    # onedim_core.py:333-339: sigma / sigma_0 > threshold (sigma_0 the largest
    # value at that bond), then [:chi]; the removed sigma_0 factors return at
    # the end of the segment (:357-358).  An absolute rule keeps abs_keep.
    sigma = scale * np.array([1.0, 1e-3, 1e-9, 1e-16])
    assert np.count_nonzero(sigma > threshold) == abs_keep
    arrs, U, V, _ = schmidt_pair(sigma, 8, 5)
    m = o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrs])
    m.left_canonise(threshold=threshold)
    assert m.bonddims() == [1, rel_keep, 1]
    expect = (U[:, :rel_keep] * sigma[:rel_keep]) @ V[:, :rel_keep].T
    got = om.mps_to_dense(m).reshape(8, 8)
    np.testing.assert_allclose(got, expect, atol=1e-13 * scale)
    # chi caps after the threshold
    m = o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrs])
    m.left_canonise(threshold=threshold, chi=2)
    assert m.bonddims() == [1, 2, 1]


def test_init_mps_random_closed_form():
    # onedim_utils.py:23-60: site i = rand draw / its largest singular value
    np.random.seed(11)
    m = o.init_mps_random(5, 2, [3, 4, 4, 2])
    np.random.seed(11)
    dims = [1, 3, 4, 4, 2, 1]
    for i, t in enumerate(m):
        raw = np.random.rand(2, dims[i], dims[i + 1])
        expect = raw / np.linalg.norm(raw.reshape(2 * dims[i], dims[i + 1]), 2)
        assert t.labels == ["phys", "left", "right"]
        np.testing.assert_allclose(t.data, expect, rtol=1e-12, atol=1e-14)
    assert m.bonddims() == dims


def test_schmidt_chain_has_the_given_spectrum():
    # the builder of the decaying-spectrum tests: the middle bond's Schmidt
    # values of the dense state are exactly sigma
    sigma = decaying_sigma(16, 3.0)
    arrs = schmidt_chain(10, 2, sigma, 1)
    m = o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrs])
    v = om.mps_to_dense(m).reshape(2 ** 5, 2 ** 5)
    np.testing.assert_allclose(np.linalg.svd(v, compute_uv=False)[:16], sigma, rtol=1e-12)


def test_svd_compress_of_a_zero_state_raises_like_the_reference():
    # onedim_core.py:476-480 divides by the zero norm; the right sweep's SVD of
    # the NaN site then fails in both LAPACK drivers (tensor.py:914-929)
    arrs = [np.zeros((2, 1 if i == 0 else 3, 1 if i == 3 else 3)) for i in range(4)]
    m = o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrs])
    with pytest.warns(Warning):
        with pytest.raises(ValueError):
            m.svd_compress(chi=2)
