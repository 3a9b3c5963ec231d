"""Host-side logic of the GPU path that needs no GPU: the structured sweep's
bond-rank rule and its resolution floor."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tncontract_amd.onedim.structured import _EPS_K, bond_rank  # noqa: E402


def test_bond_rank_follows_the_reference_rule_above_the_floor():
    # onedim_core.py:333-339 restated on sigma^2: keep sigma/sigma_0 >
    # threshold, then at most chi; at most avail
    lam = np.exp(-np.arange(64) / 2.0)          # sigma_k = exp(-k/4)
    k, rel, f = bond_rank(lam, 64, 64, None, 2e-6, _EPS_K["fp64"])
    assert k == int(np.count_nonzero(np.exp(-np.arange(64) / 4.0) > 2e-6)) == 53
    np.testing.assert_allclose(rel, np.exp(-np.arange(64) / 4.0), rtol=1e-14)
    assert f < 2e-6
    assert bond_rank(lam, 64, 64, 12, 1e-15, _EPS_K["fp64"])[0] == 12
    assert bond_rank(lam, 64, 10, 12, 1e-15, _EPS_K["fp64"])[0] == 10   # avail caps


def test_bond_rank_refuses_decisions_below_the_floor():
    lam = np.exp(-np.arange(64) / 2.0)
    # FP32-formed K: floor ~1e-2, so a cut at 1e-3 or a threshold of 2e-6 is unresolved
    assert bond_rank(lam, 64, 64, 28, 1e-15, _EPS_K["fp32"])[0] is None
    assert bond_rank(lam, 64, 64, None, 2e-6, _EPS_K["fp32"])[0] is None
    assert bond_rank(lam, 64, 64, 12, 1e-15, _EPS_K["fp32"])[0] == 12
    # rank-deficient Gram (exact zeros plus rounding noise): threshold-bound
    noisy = np.concatenate([np.ones(5), 1e-17 * np.ones(11)])
    assert bond_rank(noisy, 16, 16, None, 1e-15, _EPS_K["fp64"])[0] is None
    # zero, negative or non-finite leading eigenvalue: unresolved
    assert bond_rank(np.zeros(4), 4, 4, 2, 1e-15, _EPS_K["fp64"])[0] is None
    assert bond_rank(np.array([np.nan, 1.0]), 2, 2, 1, 1e-15, _EPS_K["fp64"])[0] is None


def test_floor_matches_the_eigensolver_shift():
    # f = sqrt(4 delta / lambda_0), delta = 8 sqrt(n) eps tr K
    lam = np.array([4.0, 1.0, 0.25, 0.0])
    _, _, f = bond_rank(lam, 4, 4, 2, 1e-15, _EPS_K["fp32"])
    delta = 8 * 2.0 * 2.0 ** -23 * 5.25
    assert abs(f - np.sqrt(4 * delta / 4.0)) < 1e-15
