"""Host-side logic of the GPU path that needs no GPU: the structured sweep's
bond-rank rule and its resolution floor, and bench.py's one-GPU contract."""
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
    # avail caps: the rank bound 10, K's eigenvalues past it rounding-level
    capped = np.concatenate([lam[:10], 1e-18 * np.ones(54)])
    assert bond_rank(capped, 64, 10, 12, 1e-15, _EPS_K["fp64"])[0] == 10


def test_bond_rank_refuses_decisions_below_the_floor():
    lam = np.exp(-np.arange(64) / 2.0)
    # an FP32-formed K whose asymmetry puts the floor near 1e-2: a cut at
    # 1e-3 or a threshold of 2e-6 is unresolved, a cut at 6e-2 is not
    assert bond_rank(lam, 64, 64, 28, 1e-15, _EPS_K["fp32"], asym=5e-5)[0] is None
    assert bond_rank(lam, 64, 64, None, 2e-6, _EPS_K["fp32"], asym=5e-5)[0] is None
    assert bond_rank(lam, 64, 64, 12, 1e-15, _EPS_K["fp32"], asym=5e-5)[0] == 12
    # rank-deficient Gram (exact zeros plus rounding noise) with a rank bound
    # above the rank: threshold-bound, unresolved
    noisy = np.concatenate([np.ones(5), 1e-17 * np.ones(11)])
    assert bond_rank(noisy, 16, 16, None, 1e-15, _EPS_K["fp64"])[0] is None
    # ... but resolved when the rank bound (avail) is the rank: the values
    # past it are K's structural null space and set the noise level
    assert bond_rank(noisy, 16, 5, None, 1e-15, _EPS_K["fp64"])[0] == 5
    # zero, negative or non-finite leading eigenvalue: unresolved
    assert bond_rank(np.zeros(4), 4, 4, 2, 1e-15, _EPS_K["fp64"])[0] is None
    assert bond_rank(np.array([np.nan, 1.0]), 2, 2, 1, 1e-15, _EPS_K["fp64"])[0] is None


def test_floor_levels():
    # f = sqrt(FLOOR_C * level / lambda_0), level = max(ASYM_C * 2 asym /
    # sqrt(n), eps sqrt(n) lambda_0, |eigenvalues past avail|)
    from tncontract_amd.onedim.structured import ASYM_C, FLOOR_C
    lam = np.array([4.0, 1.0, 0.25, 0.0])
    _, _, f = bond_rank(lam, 4, 4, 2, 1e-15, _EPS_K["fp32"], asym=1e-3)
    assert abs(f - np.sqrt(FLOOR_C * ASYM_C * 2.0 * 1e-3 / 2.0 / 4.0)) < 1e-15
    _, _, f = bond_rank(lam, 4, 4, 2, 1e-15, _EPS_K["fp32"])
    assert abs(f - np.sqrt(FLOOR_C * 2.0 ** -23 * 2.0 * 4.0 / 4.0)) < 1e-15
    _, _, f = bond_rank(lam, 4, 2, 2, 1e-15, _EPS_K["fp32"])
    assert abs(f - np.sqrt(FLOOR_C * 0.25 / 4.0)) < 1e-15


def test_learned_features_and_samples():
    from tncontract_amd import learned as lr
    rel = np.exp(-np.arange(40) / 5.0)
    f = lr.features(rel, 512, 0.25)
    assert f.shape == (34,) and f.dtype == np.float32
    np.testing.assert_allclose(f[:32], rel[:32], rtol=1e-6)
    assert abs(f[32] - np.log(512)) < 1e-6 and abs(f[33] - 0.25) < 1e-7
    short = lr.features(rel[:5], 16, 0.5)
    assert np.all(short[5:32] == 0)
    # the previous sweep's spectrum at bond b pairs with this sweep's kept rank at b
    X, T = lr.training_samples({1: rel, 2: rel[:10]}, {1: 100, 2: 7, 3: 1}, 128, 4)
    assert X.shape == (2, 34) and np.allclose(T, [100 / 128, 7 / 128])
    assert abs(X[0, 33] - 0.25) < 1e-7 and abs(X[1, 33] - 0.5) < 1e-7


def test_right_sweep_bond_numbering():
    # svd_compress's right sweep runs on the reversed chain: its SVD at
    # reversed position i splits original sites N-2-i | N-1-i, i.e. bond N-1-i
    from tncontract_amd.learned import _bond_of
    assert [_bond_of(i, 6) for i in range(5)] == [5, 4, 3, 2, 1]


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says
    so, before any GPU work."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=root)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr
    assert r.stdout == ""


def test_stride_fusion_predicate():
    # decides whether a contraction operand is copied into a dense
    # (free, contracted) order before the MFMA product
    import torch
    from tncontract_amd.tensor import _fuses
    C = torch.zeros(4, 3, 5, 2)                 # dense (a, k, al, t)
    assert _fuses(C, [0, 1, 2, 3]) and _fuses(C, [2, 3]) and _fuses(C, [1, 2])
    assert not _fuses(C, [0, 2])                # a and al are not neighbours
    assert not _fuses(C, [3, 1])                # t before k: strides do not nest
    assert _fuses(C, [0]) and _fuses(C, [])
    P = C.permute(0, 2, 1, 3)                   # (a, al, k, t) view of the same data
    assert not _fuses(P, [0, 1])
    assert _fuses(P.contiguous(), [0, 1])
    U = torch.zeros(4, 1, 5)                    # size-1 axes do not break a fusion
    assert _fuses(U, [0, 1, 2])
