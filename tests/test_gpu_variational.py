"""Parity of tensor_lq, ladder_contract and variational_compress with the CPU
oracle, on the GPU (onedim_core.py:486-631, :1474-1478, :1491-1663;
tensor.py:1058-1109).

Bar: FP64 tensors equal to the oracle to <= 1e-11 (relative) where the
result is gauge-fixed (ladder contractions), MPS results compared through the
state error ||psi_gpu - psi_oracle||^2 / ||psi_oracle||^2 <= 1e-10, and at
sizes the oracle cannot reach through the projection identities a converged
single-site variational sweep satisfies: <w|v> = <w|w> and
|v - w|^2 = |v|^2 - |w|^2, with |v - w| never above the svd_compress start."""
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


def gpu_mps(tn, arrs, labels=("phys", "left", "right")):
    return tn.onedim.MatrixProductState([tn.Tensor(a, list(labels)) for a in arrs])


def o_mps(arrs, labels=("phys", "left", "right")):
    return o.OMPS([o.OTensor(a, list(labels)) for a in arrs])


def to_dense_gpu(mps):
    return om.mps_to_dense(o.OMPS([o.OTensor(t.numpy(), t.labels) for t in mps],
                                  mps.left_label, mps.right_label, mps.phys_label))


def state_err(a, b):
    return np.linalg.norm(a - b) ** 2 / np.linalg.norm(b) ** 2


def labelled(data, labels, order):
    return np.transpose(data, [labels.index(x) for x in order])


@pytest.mark.parametrize("shape,rows", [((3, 4, 5), ["a", "b"]), ((6, 2, 9), ["c"]),
                                        ((7, 3, 2), ["a"]), ((1, 5, 4), ["a", "c"])])
def test_tensor_lq_matches_oracle(tn, shape, rows):
    rng = np.random.default_rng(sum(shape))
    arr = rng.standard_normal(shape)
    Lg, Qg = tn.tensor_lq(tn.Tensor(arr, ["a", "b", "c"]), rows, lq_label="z")
    Lo, Qo = o.tensor_lq(o.OTensor(arr, ["a", "b", "c"]), rows, lq_label="z")
    assert Lg.labels == Lo.labels and Qg.labels == Qo.labels
    assert Lg.shape == Lo.shape and Qg.shape == Qo.shape
    # L Q reproduces the input; Q rows orthonormal; each Q row equals the
    # oracle's up to the sign the Householder convention picks
    k = Qg.shape[Qg.labels.index("zout")]
    q = labelled(Qg.numpy(), Qg.labels, ["zout"] + [x for x in Qg.labels if x != "zout"]).reshape(k, -1)
    qo = labelled(Qo.data, Qo.labels, ["zout"] + [x for x in Qo.labels if x != "zout"]).reshape(k, -1)
    np.testing.assert_allclose(q @ q.T, np.eye(k), atol=1e-13)
    np.testing.assert_allclose(np.abs(q @ qo.T), np.eye(k), atol=1e-10)
    prod = tn.contract(Lg, Qg, "zin", "zout")
    np.testing.assert_allclose(labelled(prod.numpy(), prod.labels, ["a", "b", "c"]), arr, atol=1e-13)


def test_ladder_contract_matches_oracle(tn):
    from tncontract_amd.onedim import random_mps_arrays
    N = 7
    a1 = random_mps_arrays(N, 2, 4, seed=11)
    a2 = random_mps_arrays(N, 2, 5, seed=12)
    g1, g2, r1, r2 = gpu_mps(tn, a1), gpu_mps(tn, a2), o_mps(a1), o_mps(a2)
    cases = [dict(), dict(end=3), dict(start=2, end=N - 1), dict(start=1, end=4),
             dict(start=0, end=0), dict(start=N - 1, end=N - 1)]
    for kw in cases:
        for inter in (False, True):
            g = tn.onedim.ladder_contract(g1, g2, "phys", "phys", return_intermediate_contractions=inter,
                                          **kw)
            r = o.ladder_contract(r1, r2, "phys", "phys", return_intermediate_contractions=inter, **kw)
            gl, rl = (g, r) if inter else ([g], [r])
            assert len(gl) == len(rl), kw
            for x, y in zip(gl, rl):
                assert x.labels == y.labels, kw
                np.testing.assert_allclose(x.numpy(), y.data, rtol=1e-11,
                                           atol=1e-12 * np.abs(y.data).max(), err_msg=str(kw))
    with pytest.raises(ValueError):
        tn.onedim.ladder_contract(g1, g2, "phys", "phys", start=3, end=2)


def xxz_target(tn, N, chi, seed):
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo
    arrs = random_mps_arrays(N, 2, chi, seed=seed)
    g = tn.onedim.contract_mps_mpo(gpu_mps(tn, arrs), xxz_mpo(N, 1.0, 0.7, 0.3))
    r = o.contract_mps_mpo(o_mps(arrs), om.xxz_mpo(N, 1.0, 0.7, 0.3))
    return g, r


@pytest.mark.parametrize("N,chi_in,chi", [(6, 3, 8), (9, 4, 2), (10, 6, 5), (12, 8, 16)])
def test_variational_compress_matches_oracle(tn, N, chi_in, chi):
    g, r = xxz_target(tn, N, chi_in, seed=N + chi)
    outg = g.variational_compress(chi, max_iter=60, tolerance=1e-8)
    outr = r.variational_compress(chi, max_iter=60, tolerance=1e-8)
    assert (outg.left_label, outg.right_label, outg.phys_label) == \
        (outr.left_label, outr.right_label, outr.phys_label)
    assert outg.bonddims() == outr.bonddims()
    assert [t.labels for t in outg] == [t.labels for t in outr]
    assert state_err(to_dense_gpu(outg), om.mps_to_dense(outr)) < 1e-10
    assert outg.check_canonical_form(threshold=1e-10, print_output=False)[0] == N - 1
    # the target is left as it was
    assert state_err(to_dense_gpu(g), om.mps_to_dense(r)) < 1e-24


def test_variational_compress_normalise_initial_guess_functional(tn):
    from tncontract_amd.onedim import random_mps_arrays
    g, r = xxz_target(tn, 8, 4, seed=5)
    ga = random_mps_arrays(8, 2, 3, seed=6)
    outg = g.variational_compress(3, max_iter=60, tolerance=1e-8, normalise=True,
                                  initial_guess=gpu_mps(tn, ga))
    outr = r.variational_compress(3, max_iter=60, tolerance=1e-8, normalise=True,
                                  initial_guess=o_mps(ga))
    assert state_err(to_dense_gpu(outg), om.mps_to_dense(outr)) < 1e-10
    np.testing.assert_allclose(outg.norm(), 1.0, rtol=1e-12)
    f = tn.onedim.variational_compress_mps(g, 3, max_iter=60, tolerance=1e-8)
    fr = o.variational_compress_mps(r, 3, max_iter=60, tolerance=1e-8)
    assert state_err(to_dense_gpu(f), om.mps_to_dense(fr)) < 1e-10
    with pytest.raises(RuntimeError):
        g.variational_compress(2, max_iter=1, tolerance=1e-30)


def test_variational_compress_large_projection_identities(tn):
    """A 24-site, bond-320 target compressed to chi=48: beyond the dense
    oracle, so checked through the identities of the final projection."""
    from tncontract_amd.onedim import inner_product_mps, frob_distance_squared
    g, _ = xxz_target(tn, 24, 64, seed=3)
    v2 = inner_product_mps(g, g)
    s = g.copy()
    s.svd_compress(chi=48)
    w = g.variational_compress(48, max_iter=100, tolerance=1e-6)
    assert max(w.bonddims()) <= 48
    assert w.check_canonical_form(threshold=1e-10, print_output=False)[0] == 23
    wv = inner_product_mps(w, g)
    ww = inner_product_mps(w, w)
    np.testing.assert_allclose(wv, ww, rtol=1e-9)
    d_var = frob_distance_squared(w, g)
    np.testing.assert_allclose(d_var, v2 - ww, rtol=1e-6, atol=1e-12 * v2)
    assert d_var <= frob_distance_squared(s, g) * (1 + 1e-9)
