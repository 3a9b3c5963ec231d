"""Parity of the drop-in API (tncontract_amd) with the CPU oracle, on the GPU.

Bar: FP64 results equal to the oracle to <= 1e-10 (relative), and every MPS
result compared through gauge-invariant quantities: the state error
||psi_gpu - psi_oracle||^2 / ||psi_oracle||^2, norms, overlaps, singular
values and <Sz_i>."""
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


def both(tn, arr, labels):
    return tn.Tensor(arr, labels), o.OTensor(arr, labels)


def test_contract_and_shorthand(tn):
    rng = np.random.default_rng(0)
    A, oA = both(tn, rng.standard_normal((2, 2)), ["spam", "eggs"])
    B, oB = both(tn, rng.standard_normal((2, 3, 2, 4)), [])
    C = tn.contract(A, B, "spam", "i2")
    oC = o.contract(oA, oB, "spam", "i2")
    assert C.labels == oC.labels == ["eggs", "i0", "i1", "i3"]
    np.testing.assert_allclose(C.numpy(), oC.data, rtol=1e-13)
    D = A["spam", "eggs"] * B["i0", "i2"]
    oD = o.contract(oA, oB, ["spam", "eggs"], ["i0", "i2"])
    assert D.labels == oD.labels
    np.testing.assert_allclose(D.numpy(), oD.data, rtol=1e-13)


def test_contract_duplicate_labels_slices_and_errors(tn):
    rng = np.random.default_rng(1)
    A, oA = both(tn, rng.standard_normal((3, 4, 3)), ["x", "y", "x"])
    B, oB = both(tn, rng.standard_normal((3, 5, 3)), ["x", "z", "x"])
    np.testing.assert_allclose(tn.contract(A, B, "x", "x").numpy(), o.contract(oA, oB, "x", "x").data,
                               rtol=1e-13)
    C = tn.contract(A, B, "x", "x", index_slice1=[-1], index_slice2=[0])
    oC = o.contract(oA, oB, "x", "x", index_slice1=[-1], index_slice2=[0])
    assert C.labels == oC.labels
    np.testing.assert_allclose(C.numpy(), oC.data, rtol=1e-13)
    with pytest.raises(ValueError):
        tn.contract(A, B, "y", "z")


def test_tensordot_numpy_semantics(tn):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((4, 5, 6))
    b = rng.standard_normal((6, 5, 3))
    np.testing.assert_allclose(tn.tensordot(a, b, ([1, 2], [1, 0])).cpu().numpy(),
                               np.tensordot(a, b, ([1, 2], [1, 0])), rtol=1e-13)
    np.testing.assert_allclose(tn.tensordot(a, b[:, :, :1].reshape(6, 5), ([2, 1], [0, 1])).cpu().numpy(),
                               np.tensordot(a, b[:, :, :1].reshape(6, 5), ([2, 1], [0, 1])), rtol=1e-13)


def test_axis_ops_and_consolidate(tn):
    rng = np.random.default_rng(3)
    arr = rng.standard_normal((2, 3, 4, 5))
    t, ot = both(tn, arr, ["z", "a", "z", "b"])
    t.consolidate_indices()
    ot.consolidate_indices()
    assert t.labels == ot.labels and t.shape == ot.shape
    np.testing.assert_array_equal(t.numpy(), ot.data)
    t, ot = both(tn, rng.standard_normal((2, 3, 4, 5, 6)), ["a", "b", "c", "b", "d"])
    t.move_indices(["d", "b", "c"], 0)
    ot.move_indices(["d", "b", "c"], 0)
    assert t.labels == ot.labels
    np.testing.assert_array_equal(t.numpy(), ot.data)


def test_trace_and_con(tn):
    rng = np.random.default_rng(4)
    A, oA = both(tn, rng.standard_normal((3, 2, 4)), ["a", "b", "c"])
    B, oB = both(tn, rng.standard_normal((3, 4)), ["d", "e"])
    C, oC = both(tn, rng.standard_normal((5, 5, 2)), ["f", "g", "h"])
    r = tn.con(A, B, C, ("a", "d"), ("c", "e"), ("f", "g"), ("h", "b"))
    orr = o.con(oA, oB, oC, ("a", "d"), ("c", "e"), ("f", "g"), ("h", "b"))
    np.testing.assert_allclose(r.numpy(), orr.data, rtol=1e-12)
    N = 30
    a = rng.standard_normal((2, 2)) / 1.5
    ts = [tn.Tensor(a, ["left", "right"]).suf(str(i)) for i in range(N)]
    pairs = [("right" + str(j), "left" + str(j + 1)) for j in range(N - 1)]
    r = tn.con(ts, pairs, ("right" + str(N - 1), "left0"))
    np.testing.assert_allclose(r.numpy(), np.trace(np.linalg.matrix_power(a, N)), rtol=1e-10)


def test_tensor_svd_qr_truncated(tn):
    rng = np.random.default_rng(5)
    arr = rng.standard_normal((6, 7, 8))
    t, ot = both(tn, arr, ["i0", "i1", "i2"])
    U, S, V = tn.tensor_svd(t, ["i0", "i2"])
    oU, oS, oV = o.tensor_svd(ot, ["i0", "i2"])
    assert U.labels == oU.labels and U.shape == oU.shape and V.labels == oV.labels
    np.testing.assert_allclose(np.diag(S.numpy()), np.diag(oS.data), rtol=1e-12)
    rec = tn.contract(U, tn.contract(S, V, "svd_in", "svd_out"), "svd_in", "svd_out")
    assert tn.distance(rec, t) < 1e-12 * np.linalg.norm(arr)
    Q, R = tn.tensor_qr(t, ["i0", "i2"])
    oQ, oR = o.tensor_qr(ot, ["i0", "i2"])
    assert Q.labels == oQ.labels and R.labels == oR.labels and Q.shape == oQ.shape
    np.testing.assert_allclose(R.numpy(), oR.data, atol=1e-11)
    u, _ = np.linalg.qr(rng.standard_normal((12, 6)))
    v, _ = np.linalg.qr(rng.standard_normal((10, 6)))
    s = np.array([5.0, 2.0, 1.0, 1e-3, 1e-9, 1e-12])
    m = (u * s) @ v.T
    for kw in [dict(chi=5, threshold=1e-6), dict(chi=0, threshold=1e-4, absolute=False),
               dict(chi=3, absorb_singular_values="left"), dict(chi=4, absorb_singular_values="both")]:
        g = tn.truncated_svd(tn.Tensor(m, ["r", "c"]), ["r"], **kw)
        r = o.truncated_svd(o.OTensor(m, ["r", "c"]), ["r"], **kw)
        assert g[0].shape == r[0].shape and g[1].shape == r[1].shape
        np.testing.assert_allclose(np.sort(np.abs(g[2])), np.sort(np.abs(r[2])), rtol=1e-6, atol=1e-14)
        prod_g = tn.contract(g[0], g[1], "svd_in", "svd_out").numpy()
        prod_o = o.contract(r[0], r[1], "svd_in", "svd_out").data
        np.testing.assert_allclose(prod_g, prod_o, atol=1e-12)


def gpu_mps(tn, arrs, labels=("phys", "left", "right")):
    return tn.onedim.MatrixProductState([tn.Tensor(a, list(labels)) for a in arrs])


def o_mps(arrs, labels=("phys", "left", "right")):
    return o.OMPS([o.OTensor(a, list(labels)) for a in arrs])


def to_dense_gpu(mps):
    return om.mps_to_dense(o.OMPS([o.OTensor(t.numpy(), t.labels) for t in mps],
                                  mps.left_label, mps.right_label, mps.phys_label))


def state_err(a, b):
    return np.linalg.norm(a - b) ** 2 / np.linalg.norm(b) ** 2


@pytest.mark.parametrize("chi", [2, 5, 16])
def test_svd_compress_matches_oracle(tn, chi):
    from tncontract_amd.onedim import random_mps_arrays
    arrs = random_mps_arrays(10, 2, 8, seed=chi)
    g, r = gpu_mps(tn, arrs), o_mps(arrs)
    g.svd_compress(chi=chi, threshold=1e-10)
    r.svd_compress(chi=chi, threshold=1e-10)
    assert [t.labels for t in g] == [t.labels for t in r]
    assert g.bonddims() == r.bonddims()
    assert state_err(to_dense_gpu(g), om.mps_to_dense(r)) < 1e-10
    assert g.check_canonical_form(threshold=1e-10, print_output=False)[1] == 0


def test_left_canonise_svd_and_qr(tn):
    from tncontract_amd.onedim import random_mps_arrays
    arrs = random_mps_arrays(8, 3, 6, seed=7)
    for kw in [dict(qr_decomposition=True), dict(), dict(normalise=True), dict(chi=3, threshold=1e-8)]:
        g, r = gpu_mps(tn, arrs), o_mps(arrs)
        g.left_canonise(**kw)
        r.left_canonise(**kw)
        assert [t.labels for t in g] == [t.labels for t in r]
        assert g.bonddims() == r.bonddims()
        assert state_err(to_dense_gpu(g), om.mps_to_dense(r)) < 1e-10
        np.testing.assert_allclose(g.norm(), r.norm(), rtol=1e-10)


def test_contract_mps_mpo_and_compress_config_small(tn):
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo
    N = 8
    arrs = random_mps_arrays(N, 2, 6, seed=1)
    g, r = gpu_mps(tn, arrs), o_mps(arrs)
    gm = tn.onedim.contract_mps_mpo(g, xxz_mpo(N))
    rm = o.contract_mps_mpo(r, om.xxz_mpo(N))
    assert [t.labels for t in gm] == [t.labels for t in rm]
    assert [t.shape for t in gm] == [t.shape for t in rm]
    for a, b in zip(gm, rm):
        np.testing.assert_allclose(a.numpy(), b.data, rtol=1e-13, atol=1e-15)
    gm.svd_compress(chi=6)
    rm.svd_compress(chi=6)
    assert state_err(to_dense_gpu(gm), om.mps_to_dense(rm)) < 1e-10


def test_inner_product_distance_expvals(tn):
    from tncontract_amd.onedim import random_mps_arrays
    a1 = random_mps_arrays(7, 2, 4, seed=2)
    a2 = random_mps_arrays(7, 2, 4, seed=3)
    g1, g2 = gpu_mps(tn, a1), gpu_mps(tn, a2)
    r1, r2 = o_mps(a1), o_mps(a2)
    np.testing.assert_allclose(tn.onedim.inner_product_mps(g1, g2), o.inner_product_mps(r1, r2), rtol=1e-11)
    np.testing.assert_allclose(tn.onedim.frob_distance_squared(g1, g2), o.frob_distance_squared(r1, r2),
                               rtol=1e-10)
    op = tn.Tensor(om.SZ, ["out", "in"])
    ev = tn.onedim.expvals_mps(g1, op)
    oev = o.expvals_mps(r1, o.OTensor(om.SZ, ["out", "in"]))
    np.testing.assert_allclose(ev.real, oev.real, atol=1e-11)


def test_reference_fixture_invariants_gpu(tn, fixture_mps_arrays):
    # tncontract/tests/test_canonical_form_mps_conversion.py invariants
    arrays, meta = fixture_mps_arrays
    psi = tn.onedim.MatrixProductState([tn.Tensor(a, lab) for a, lab in zip(arrays, meta["labels"])],
                                       meta["left_label"], meta["right_label"], meta["phys_label"])
    n2 = float(tn.onedim.inner_product_mps(psi, psi))
    np.testing.assert_allclose(n2, 0.4668553069796, atol=1e-12)
    psi.svd_compress(threshold=1e-12, normalise=False)
    np.testing.assert_almost_equal(psi.norm(), np.sqrt(n2), decimal=10)
    assert psi.check_canonical_form(threshold=1e-10, print_output=False)[1] == 0
    psi.left_canonise()
    assert psi.check_canonical_form(threshold=1e-10, print_output=False)[0] == len(psi) - 1
    psi.svd_compress(threshold=1e-12, normalise=True)
    np.testing.assert_almost_equal(psi.norm(), 1.0, decimal=10)


def test_zero_state(tn):
    arrs = [np.zeros((2, 1 if i == 0 else 3, 1 if i == 3 else 3)) for i in range(4)]
    g = gpu_mps(tn, arrs)
    g.left_canonise()
    assert all(t.shape == (2, 1, 1) for t in g)


@pytest.mark.parametrize("m,n", [(30000, 40), (24577, 9), (40000, 70)])
def test_tall_qr_and_svd_beyond_register_panel(tn, m, n):
    """QR / SVD of matrices taller than one register panel (24576 rows):
    the grid-parallel column path, same Householder sign convention."""
    rng = np.random.default_rng(m + n)
    a = rng.standard_normal((m, n))
    Q, R = tn.tensor_qr(tn.Tensor(a, ["r", "c"]), ["r"])
    Qo, Ro = o.tensor_qr(o.OTensor(a, ["r", "c"]), ["r"])
    assert Q.labels == Qo.labels and R.labels == Ro.labels
    np.testing.assert_allclose(R.numpy(), Ro.data, rtol=1e-9, atol=1e-11 * np.abs(Ro.data).max())
    np.testing.assert_allclose(Q.numpy(), Qo.data, atol=1e-11)
    U, S, V = tn.tensor_svd(tn.Tensor(a, ["r", "c"]), ["r"])
    s_ref = np.linalg.svd(a, compute_uv=False)
    np.testing.assert_allclose(np.diag(S.numpy()), s_ref, rtol=1e-11)
    rec = U.numpy() @ S.numpy() @ V.numpy()
    np.testing.assert_allclose(rec, a, atol=1e-11 * np.abs(a).max())


def test_tensor_to_matrix_and_back(tn):
    # tensor.py:805-827: bit-exact data movement, same shapes and labels
    rng = np.random.default_rng(21)
    a = rng.standard_normal((2, 3, 4, 5))
    g, r = both(tn, a, ["a", "b", "c", "d"])
    for rows in (["c", "a"], [], ["d", "c", "b", "a"], ["b"]):
        Mg = tn.tensor_to_matrix(g, rows)
        Mr = o.tensor_to_matrix(r, rows)
        assert tuple(Mg.shape) == Mr.shape
        assert np.array_equal(Mg.cpu().numpy(), Mr)
    assert g.labels == ["a", "b", "c", "d"]
    m = rng.standard_normal((6, 20))
    tg = tn.matrix_to_tensor(m, (2, 3, 4, 5))
    tr = o.matrix_to_tensor(m, (2, 3, 4, 5))
    assert tg.labels == tr.labels == ["i0", "i1", "i2", "i3"]
    assert np.array_equal(tg.numpy(), tr.data)
    tg = tn.matrix_to_tensor(tn.tensor_to_matrix(g, ["c", "a"]), (4, 2, 3, 5), ["c", "a", "b", "d"])
    assert tg.labels == ["c", "a", "b", "d"]
    assert np.array_equal(tg.numpy(), np.transpose(a, (2, 0, 1, 3)))


@pytest.mark.parametrize("scale,threshold,rel_keep", [(1e-3, 1e-10, 3), (1e3, 1e-14, 3)])
def test_left_canonise_threshold_is_relative_gpu(tn, scale, threshold, rel_keep):
    # onedim_core.py:333-339: sigma / sigma_0 > threshold; an absolute rule
    # would keep 2 (scale 1e-3) or 4 (scale 1e3) of these Schmidt values
    from cases import schmidt_pair
    sigma = scale * np.array([1.0, 1e-3, 1e-9, 1e-16])
    arrs, U, V, _ = schmidt_pair(sigma, 8, 5)
    g, r = gpu_mps(tn, arrs), o_mps(arrs)
    g.left_canonise(threshold=threshold)
    r.left_canonise(threshold=threshold)
    assert g.bonddims() == r.bonddims() == [1, rel_keep, 1]
    expect = ((U[:, :rel_keep] * sigma[:rel_keep]) @ V[:, :rel_keep].T).reshape(-1)
    assert state_err(to_dense_gpu(g), expect) < 1e-20
    assert state_err(to_dense_gpu(g), om.mps_to_dense(r)) < 1e-20


def test_init_mps_random_matches_oracle(tn):
    # onedim_utils.py:23-60 with the same np.random stream
    np.random.seed(3)
    g = tn.onedim.init_mps_random(6, 2, [4, 5, 5, 3, 2])
    np.random.seed(3)
    r = o.init_mps_random(6, 2, [4, 5, 5, 3, 2])
    assert g.bonddims() == r.bonddims() == [1, 4, 5, 5, 3, 2, 1]
    for a, b in zip(g, r):
        assert a.labels == b.labels
        np.testing.assert_allclose(a.numpy(), b.data, rtol=1e-12, atol=1e-14)
        s = np.linalg.svd(a.numpy().reshape(-1, a.shape[2]), compute_uv=False)
        assert abs(s[0] - 1.0) < 1e-12


def test_svd_nonconvergence_is_retried(tn):
    # tensor.py:914-921: a failed SVD is retried (there with gesvd) after a
    # warning.  Forced from the host with a one-sweep limit.
    from tncontract_amd import tensor as tsr
    rng = np.random.default_rng(5)
    a = rng.standard_normal((60, 40))
    with pytest.warns(UserWarning, match="did not converge"):
        u, s, vt = tsr._svd_device(torch.from_numpy(a).cuda(), 1, max_sweeps=1)
    np.testing.assert_allclose(s.cpu().numpy(), np.linalg.svd(a, compute_uv=False), rtol=1e-12)
    rec = (u.cpu().numpy() * s.cpu().numpy()) @ vt.cpu().numpy()
    np.testing.assert_allclose(rec, a, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_svd_of_nonfinite_matrix_raises_like_the_reference(tn, capsys, bad):
    # tensor.py:922-929: the diagnostic line, then ValueError
    a = np.ones((6, 5))
    a[2, 3] = bad
    with pytest.raises(ValueError):
        tn.tensor_svd(tn.Tensor(a, ["r", "c"]), ["r"])
    out = capsys.readouterr().out
    assert ("Matrix contains inf's: %s. Matrix contains nan's: %s"
            % (bool(np.isinf(bad)), bool(np.isnan(bad)))) in out
