"""The benchmark configurations at their stated sizes, on the GPU, against the
NumPy FP64 oracle on the same seeded inputs.

Bars (every error is against the oracle):
  * plain FP64 path: state error ||psi - psi_oracle||^2 / ||psi_oracle||^2
    <= 1e-10, identical bond dimensions, labels and axis orders;
  * structured sweep with FP32 / bf16x3 contractions: state error <= 1e-8;
  * local expectation values: |d<O>| <= 2 ||O|| sqrt(err) bounds the change
    of <O> by a state error err, so with ||Sz|| = 1/2 the bars are 1e-5
    (plain, from 1e-10) and 1e-4 (mixed, from 1e-8).
Overlaps and state errors are FP64 transfer-matrix contractions on the host
(oracle.inner_product_mps) of the device result copied back."""
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


def to_oracle(g):
    return o.OMPS([o.OTensor(t.numpy().astype(np.float64), t.labels) for t in g],
                  g.left_label, g.right_label, g.phys_label)


def rel_dist2(a, b):
    """||a - b||^2 / ||b||^2 of two oracle MPSs (FP64 overlaps)."""
    nb = o.inner_product_mps(b, b)
    return float((o.inner_product_mps(a, a) + nb - 2 * o.inner_product_mps(a, b)) / nb)


def gpu_state(tn, arrs):
    return tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"]) for a in arrs])


def oracle_state(arrs):
    return o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrs])


def test_config1_left_canonise_then_svd_compress(tn):
    """Config 1: N=20, d=2, chi=32, i.i.d. N(0,1) normalised; left_canonise()
    then svd_compress(chi=16, threshold=1e-10); overlap with the uncompressed
    state and truncation error equal to the oracle's."""
    from tncontract_amd.onedim import random_mps_arrays
    arrs = random_mps_arrays(20, 2, 32, seed=0)
    g, r = gpu_state(tn, arrs), oracle_state(arrs)
    g.left_canonise()
    r.left_canonise()
    assert g.bonddims() == r.bonddims()
    g_in, r_in = to_oracle(g), r.copy()
    g.svd_compress(chi=16, threshold=1e-10)
    r.svd_compress(chi=16, threshold=1e-10)
    assert g.bonddims() == r.bonddims() and max(r.bonddims()) == 16
    assert [t.labels for t in g] == [t.labels for t in r]
    go = to_oracle(g)
    assert rel_dist2(go, r) < 1e-10
    ov_g, ov_r = o.inner_product_mps(g_in, go), o.inner_product_mps(r_in, r)
    assert abs(ov_g - ov_r) <= 1e-10 * abs(ov_r)
    eps_g, eps_r = rel_dist2(go, g_in), rel_dist2(r, r_in)
    assert abs(eps_g - eps_r) <= 1e-10 * abs(eps_r) + 1e-14, (eps_g, eps_r)
    assert g.check_canonical_form(threshold=1e-10, print_output=False)[1] == 0


@pytest.fixture(scope="module")
def config2(tn):
    """Config 2 inputs and the oracle result: XXZ MPO (D=5, Jz=1, Jxy=1,
    hz=0.5) on a seeded random MPS N=64, d=2, chi=128, compressed from the
    fused bond 640 back to 128; <Sz_i> of the result."""
    from tncontract_amd.onedim import random_mps_arrays
    N, chi = 64, 128
    arrs = random_mps_arrays(N, 2, chi, seed=2)
    r = o.contract_mps_mpo(oracle_state(arrs), om.xxz_mpo(N))
    r.svd_compress(chi=chi, threshold=1e-15)
    # expvals_mps moves the orthogonality centre in place (onedim_utils.py:
    # 223-228): a copy keeps r as svd_compress left it
    sz = o.expvals_mps(r.copy(), o.OTensor(om.SZ, ["out", "in"])).real
    return arrs, r, sz


def _sz_gpu(tn, g):
    return tn.onedim.expvals_mps(g, tn.Tensor(om.SZ, ["out", "in"])).real


def test_config2_plain_fp64(tn, config2):
    from tncontract_amd.onedim import xxz_mpo
    arrs, r, sz = config2
    g = tn.onedim.contract_mps_mpo(gpu_state(tn, arrs), xxz_mpo(64))
    g.svd_compress(chi=128, threshold=1e-15)
    assert g.bonddims() == r.bonddims() and max(g.bonddims()) == 128
    assert [t.labels for t in g] == [t.labels for t in r]
    assert rel_dist2(to_oracle(g), r) < 1e-10
    assert np.max(np.abs(_sz_gpu(tn, g) - sz)) < 1e-5


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_config2_mixed_precision(tn, config2, prec):
    from tncontract_amd.onedim import xxz_mpo
    from tncontract_amd.onedim.structured import apply_mpo_compress, SweepStats
    arrs, r, sz = config2
    st = SweepStats()
    g = apply_mpo_compress(gpu_state(tn, arrs), xxz_mpo(64), chi=128, threshold=1e-15,
                           precision=prec, stats=st)
    assert st.fallback is None
    assert g.bonddims() == r.bonddims()
    assert [t.labels for t in g] == [t.labels for t in r]
    err = rel_dist2(to_oracle(g), r)
    assert err < 1e-8, err
    assert g.check_canonical_form(threshold=1e-10, print_output=False)[1] == 0
    assert np.max(np.abs(_sz_gpu(tn, g) - sz)) < 1e-4


def test_headline_shapes_chi512(tn):
    """The chi=512 workload's per-site shapes on a 22-site chain: bulk sites
    reach the 5120 x 2560 QR and the 1024 x 2560 SVD of the plain path and the
    n = 1024 eigensolve of the structured sweep.  Plain FP64 and the FP32
    structured sweep against the oracle."""
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo
    from tncontract_amd.onedim.structured import apply_mpo_compress, SweepStats
    N, chi = 22, 512
    arrs = random_mps_arrays(N, 2, chi, seed=0)
    r = o.contract_mps_mpo(oracle_state(arrs), om.xxz_mpo(N))
    r.svd_compress(chi=chi, threshold=1e-15)
    assert max(r.bonddims()) == chi
    g = tn.onedim.contract_mps_mpo(gpu_state(tn, arrs), xxz_mpo(N))
    g.svd_compress(chi=chi, threshold=1e-15)
    assert g.bonddims() == r.bonddims()
    assert rel_dist2(to_oracle(g), r) < 1e-10
    st = SweepStats()
    f = apply_mpo_compress(gpu_state(tn, arrs), xxz_mpo(N), chi=chi, threshold=1e-15,
                           precision="fp32", stats=st)
    assert st.fallback is None
    assert max(st.avail.values()) == 1024       # the n = 1024 eigensolve ran
    assert f.bonddims() == r.bonddims()
    err = rel_dist2(to_oracle(f), r)
    assert err < 1e-8, err


def test_config5_shapes_bf16x3(tn):
    """Config 5's per-site shapes (XXZ D=5 on chi=1024, fused bond 5120,
    compressed back to 1024) on a 24-site chain with bf16x3 contractions:
    the structured sweep's result is right-canonical to 1e-10 and within
    1e-8 (state distance) of the FP32-contraction sweep; its peak device
    memory is printed.  (The 128-site chain itself
    is the benchmark's workload; the oracle is not run at this size.)"""
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo
    from tncontract_amd.onedim.structured import apply_mpo_compress, SweepStats
    N, chi = 24, 1024
    arrs = random_mps_arrays(N, 2, chi, seed=5)
    psi = gpu_state(tn, arrs)
    torch.cuda.reset_peak_memory_stats()
    st = SweepStats()
    b = apply_mpo_compress(psi, xxz_mpo(N), chi=chi, threshold=1e-15, precision="bf16x3", stats=st)
    peak = torch.cuda.max_memory_allocated()
    assert st.fallback is None
    assert max(st.avail.values()) == 2 * chi          # the n = 2048 eigensolve ran
    assert max(b.bonddims()) == chi
    assert b.check_canonical_form(threshold=1e-10, print_output=False)[1] == 0
    f = apply_mpo_compress(psi, xxz_mpo(N), chi=chi, threshold=1e-15, precision="fp32")
    assert f.bonddims() == b.bonddims()
    ip = tn.onedim.inner_product_mps
    nf = float(ip(f, f))
    dist = (float(ip(b, b)) + nf - 2 * float(ip(b, f))) / nf
    assert abs(dist) < 1e-8, dist
    print("config 5 shapes: peak device memory %.2f GB" % (peak / 1e9))
