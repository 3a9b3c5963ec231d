"""The structured apply+compress sweep reproduces contract_mps_mpo +
svd_compress (oracle) on the GPU: same bond dimensions and labels, state
error <= 1e-10 in FP64 and <= 1e-8 with FP32 / bf16x3 contractions."""
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


def oracle_result(arrs, N, chi_out, thr):
    r = o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrs])
    r = o.contract_mps_mpo(r, om.xxz_mpo(N))
    r.svd_compress(chi=chi_out, threshold=thr)
    return r


def state_error(g, r):
    gm = o.OMPS([o.OTensor(t.numpy().astype(np.float64), t.labels) for t in g],
                g.left_label, g.right_label, g.phys_label)
    n = o.inner_product_mps(r, r)
    return float((o.inner_product_mps(gm, gm) + n - 2 * o.inner_product_mps(gm, r)) / n)


@pytest.mark.parametrize("N,chi,chi_out,prec,tol", [
    (8, 4, 4, "fp64", 1e-12), (10, 8, 6, "fp64", 1e-12), (16, 16, 16, "fp64", 1e-12),
    (16, 16, 12, "fp32", 1e-8), (16, 16, 12, "bf16x3", 1e-8), (24, 32, 32, "fp32", 1e-8)])
def test_structured_matches_oracle(tn, N, chi, chi_out, prec, tol):
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo
    from tncontract_amd.onedim.structured import apply_mpo_compress, SweepStats
    arrs = random_mps_arrays(N, 2, chi, seed=N + chi)
    psi = tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"]) for a in arrs])
    st = SweepStats()
    g = apply_mpo_compress(psi, xxz_mpo(N), chi=chi_out, threshold=1e-15, precision=prec, stats=st)
    r = oracle_result(arrs, N, chi_out, 1e-15)
    assert g.bonddims() == r.bonddims()
    assert [t.labels for t in g] == [t.labels for t in r]
    assert [t.shape for t in g] == [t.shape for t in r]
    err = state_error(g, r)
    assert abs(err) < tol, err
    assert g.check_canonical_form(threshold=1e-10, print_output=False)[1] == 0


def test_structured_matches_plain_gpu_path(tn):
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo
    from tncontract_amd.onedim.structured import apply_mpo_compress
    N, chi = 20, 24
    arrs = random_mps_arrays(N, 2, chi, seed=3)
    psi = tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"]) for a in arrs])
    plain = tn.onedim.contract_mps_mpo(psi, xxz_mpo(N))
    plain.svd_compress(chi=chi)
    fast = apply_mpo_compress(psi, xxz_mpo(N), chi=chi)
    n = tn.onedim.inner_product_mps(plain, plain)
    err = (tn.onedim.inner_product_mps(fast, fast) + n - 2 * tn.onedim.inner_product_mps(fast, plain)) / n
    assert abs(float(err)) < 1e-10
