"""The reference's pickle format with GPU objects: the reference's fixtures
(tests/golden/*.dat) load into device MPS, the reference's own test of those
fixtures (tncontract/tests/test_canonical_form_mps_conversion.py) passes on
them, and device objects written by tncontract_amd.io read back bit-exactly.
Bar: the reference test's own tolerances (decimal=10)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def tn():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tncontract_amd
    tncontract_amd.native_library()
    return tncontract_amd


def load(tn, name):
    from tncontract_amd import io as tio
    return tio.load(os.path.join(GOLD, name))


@pytest.mark.parametrize("name", ["random_10site_mps.dat", "random_10site_mps_py2.dat"])
def test_reference_conversion_test_on_gpu(tn, name):
    od = tn.onedim
    psi = load(tn, name)
    assert isinstance(psi, od.MatrixProductState) and len(psi) == 10
    psi.svd_compress(threshold=1e-12, normalise=False)
    psicr = od.right_canonical_to_canonical(psi, threshold=1e-12)
    np.testing.assert_almost_equal(psi.norm(), psicr.norm(), decimal=10)
    l, r, n = psicr.check_canonical_form(threshold=1e-10, print_output=False)
    assert (len(l), len(r), len(n)) == (0, 0, 1)
    psi.left_canonise()
    psicl = od.left_canonical_to_canonical(psi, threshold=1e-12)
    np.testing.assert_almost_equal(float(od.inner_product_mps(psicr, psicl)), psi.norm() ** 2,
                                   decimal=10)
    psi.svd_compress(threshold=1e-12, normalise=True)
    psicr = od.right_canonical_to_canonical(psi, threshold=1e-12)
    l, r, n = psicr.check_canonical_form(threshold=1e-10, print_output=False)
    assert (len(l), len(r), len(n)) == (0, 0, 0)
    psi.left_canonise()
    psicl = od.left_canonical_to_canonical(psi, threshold=1e-12)
    np.testing.assert_almost_equal(float(od.inner_product_mps(psicr, psicl)), 1.0, decimal=10)
    psi.svd_compress(threshold=1e-12, normalise=True)
    psic = od.right_canonical_to_canonical(psi, threshold=1e-12)
    psir = od.canonical_to_right_canonical(psic)
    psil = od.canonical_to_left_canonical(psic)
    np.testing.assert_almost_equal(float(od.inner_product_mps(psi, psil)), 1.0, decimal=10)
    np.testing.assert_almost_equal(float(od.inner_product_mps(psi, psir)), 1.0, decimal=10)
    assert psir.check_canonical_form(threshold=1e-10, print_output=False)[1] == 0
    assert psil.check_canonical_form(threshold=1e-10, print_output=False)[0] == len(psi) - 1


def test_device_objects_round_trip(tn):
    from tncontract_amd import io as tio
    od = tn.onedim
    rng = np.random.default_rng(3)
    t = tn.Tensor(rng.standard_normal((3, 4, 5)), ["a", "b", "c"])
    t2 = tio.loads(tio.dumps(t))
    assert t2.labels == t.labels and np.array_equal(t2.numpy(), t.numpy())
    psi = load(tn, "random_10site_mps.dat")
    for obj in (psi, od.canonical_form_mps(psi), od.xxz_mpo(6)):
        back = tio.loads(tio.dumps(obj))
        assert type(back) is type(obj)
        assert [x.labels for x in back] == [x.labels for x in obj]
        assert all(np.array_equal(x.numpy(), y.numpy()) for x, y in zip(back, obj))
        for k in ("left_label", "right_label", "phys_label", "physout_label", "physin_label"):
            assert getattr(back, k, None) == getattr(obj, k, None)
    # a strided (permuted) device tensor is written in its logical order
    t.move_index("c", 0)
    t3 = tio.loads(tio.dumps(t))
    assert t3.labels == ["c", "a", "b"] and np.array_equal(t3.numpy(), t.numpy())
