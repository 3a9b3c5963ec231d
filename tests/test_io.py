"""The reference's on-disk format (pickled tncontract objects) on the CPU:
tncontract_amd.io reads the reference's own fixtures (tests/golden/*.dat,
copied from tncontract/tests/) without unpickling them, writes the same
layout, and the reference's own test of those fixtures
(tncontract/tests/test_canonical_form_mps_conversion.py) holds on the oracle.
No GPU call: only the host-side reader/writer is exercised here."""
import os
import pickle
import sys
import types

import numpy as np
import pytest

from oracle import tn_oracle as o
from tncontract_amd import io as tio

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    with open(os.path.join(GOLD, name), "rb") as f:
        return f.read()


def oracle_mps(attrs):
    return o.OMPS([o.OTensor(a, lab) for a, lab in attrs["tensors"]], attrs["left_label"],
                  attrs["right_label"], attrs["phys_label"])


@pytest.mark.parametrize("name", ["random_10site_mps.dat", "random_10site_mps_py2.dat"])
def test_reads_reference_fixtures(name):
    kind, attrs = tio.read_host(fixture(name))
    assert kind == "MatrixProductState"
    assert (attrs["left_label"], attrs["right_label"], attrs["phys_label"]) == \
        ("left", "right", "phys")
    assert len(attrs["tensors"]) == 10
    assert all(lab == ["phys", "left", "right"] for _, lab in attrs["tensors"])
    assert [a.shape[1:] for a, _ in attrs["tensors"]][::9] == [(1, 3), (3, 1)]
    z = np.load(os.path.join(GOLD, "random_10site_mps.npz"))
    arrs = [a for a, _ in attrs["tensors"]]
    if name == "random_10site_mps.dat":
        # the same bytes the fixture extractor turned into the committed npz
        assert all(np.array_equal(a, z["site%d" % i]) for i, a in enumerate(arrs))
    else:
        # the Python 2 file holds the same state to ~1e-9
        assert all(np.allclose(a, z["site%d" % i], atol=1e-8) for i, a in enumerate(arrs))


@pytest.mark.parametrize("name", ["random_10site_mps.dat", "random_10site_mps_py2.dat"])
def test_reference_conversion_test_on_oracle(name):
    """tncontract/tests/test_canonical_form_mps_conversion.py, both tests, on
    the oracle with the reference's fixture."""
    psi = oracle_mps(tio.read_host(fixture(name))[1])
    psi.svd_compress(threshold=1e-12, normalise=False)
    psicr = o.right_canonical_to_canonical(psi, threshold=1e-12)
    np.testing.assert_almost_equal(psi.norm(), psicr.norm(), decimal=10)
    l, r, n = psicr.check_canonical_form(threshold=1e-10)
    assert (len(l), len(r), len(n)) == (0, 0, 1)
    psi.left_canonise()
    psicl = o.left_canonical_to_canonical(psi, threshold=1e-12)
    np.testing.assert_almost_equal(o.inner_product_mps(psicr, psicl), psi.norm() ** 2, decimal=10)
    psi.svd_compress(threshold=1e-12, normalise=True)
    psicr = o.right_canonical_to_canonical(psi, threshold=1e-12)
    assert tuple(map(len, psicr.check_canonical_form(threshold=1e-10))) == (0, 0, 0)
    psi.left_canonise()
    psicl = o.left_canonical_to_canonical(psi, threshold=1e-12)
    np.testing.assert_almost_equal(o.inner_product_mps(psicr, psicl), 1.0, decimal=10)
    # test_right_canonical_to_canonical_and_back
    psi.svd_compress(threshold=1e-12, normalise=True)
    psic = o.right_canonical_to_canonical(psi, threshold=1e-12)
    psir = o.canonical_to_right_canonical(psic)
    psil = o.canonical_to_left_canonical(psic)
    np.testing.assert_almost_equal(o.inner_product_mps(psi, psil), 1.0, decimal=10)
    np.testing.assert_almost_equal(o.inner_product_mps(psi, psir), 1.0, decimal=10)
    assert psir.check_canonical_form(threshold=1e-10)[1] == 0
    assert psil.check_canonical_form(threshold=1e-10)[0] == len(psi) - 1


def test_write_read_round_trip():
    rng = np.random.default_rng(0)
    t = rng.standard_normal((2, 3, 4))
    b = tio.write_host("Tensor", t, ["a", "b", "c"])
    kind, arr, labels = tio.read_host(b)
    assert kind == "Tensor" and labels == ["a", "b", "c"] and np.array_equal(arr, t)
    sites = [(rng.standard_normal((2, 1 if i == 0 else 4, 1 if i == 5 else 4)), ["phys", "left", "right"])
             for i in range(6)]
    for kind, extra in (("MatrixProductState", {"phys_label": "p"}),
                        ("MatrixProductStateCanonical", {"phys_label": "p"}),
                        ("MatrixProductOperator", {"physout_label": "o", "physin_label": "i"})):
        attrs = dict(left_label="L", right_label="R", tensors=sites, **extra)
        k2, a2 = tio.read_host(tio.write_host(kind, attrs))
        assert k2 == kind
        assert {k: v for k, v in a2.items() if k != "tensors"} == \
            {k: v for k, v in attrs.items() if k != "tensors"}
        assert all(np.array_equal(x, y) and lx == ly
                   for (x, lx), (y, ly) in zip(a2["tensors"], sites))


def test_written_bytes_load_with_the_standard_unpickler(monkeypatch):
    """What this package writes is an ordinary pickle of the reference's
    classes: the standard unpickler (these are our own bytes) rebuilds it
    when stand-in modules provide the class names."""
    class Tensor:
        pass

    class MatrixProductState:
        pass

    mods = {"tncontract": types.ModuleType("tncontract"),
            "tncontract.tensor": types.ModuleType("tncontract.tensor"),
            "tncontract.onedim": types.ModuleType("tncontract.onedim"),
            "tncontract.onedim.onedim_core": types.ModuleType("tncontract.onedim.onedim_core")}
    mods["tncontract.tensor"].Tensor = Tensor
    mods["tncontract.onedim.onedim_core"].MatrixProductState = MatrixProductState
    for k, v in mods.items():
        monkeypatch.setitem(sys.modules, k, v)
    rng = np.random.default_rng(1)
    sites = [(rng.standard_normal((2, 1 if i == 0 else 3, 1 if i == 3 else 3)), ["phys", "left", "right"])
             for i in range(4)]
    b = tio.write_host("MatrixProductState", dict(left_label="left", right_label="right",
                                                  phys_label="phys", tensors=sites))
    obj = pickle.loads(b)
    assert isinstance(obj, MatrixProductState)
    assert list(vars(obj)) == ["left_label", "right_label", "data", "phys_label"]
    assert obj.data.dtype == object and obj.data.shape == (4,)
    for t, (a, lab) in zip(obj.data, sites):
        assert isinstance(t, Tensor) and list(vars(t)) == ["data", "_labels"]
        assert np.array_equal(t.data, a) and t._labels == lab


def test_rejects_other_classes_and_broken_streams():
    import collections
    with pytest.raises(ValueError, match="unsupported class"):
        tio.read_host(pickle.dumps(collections.OrderedDict(a=1), protocol=3))
    with pytest.raises(ValueError):
        tio.read_host(pickle.dumps([1, 2, 3], protocol=3))
    good = fixture("random_10site_mps.dat")
    with pytest.raises(Exception):
        tio.read_host(good[: len(good) // 2])
