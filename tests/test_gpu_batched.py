"""Batched uneven contraction + truncation (grouped GEMM + batched Jacobi)
against NumPy FP64 on the same FP32 inputs.  Inputs are FP32 and the
contraction accumulates in FP32, so singular values are compared to 1e-4
relative above 1e-5 * sigma_0 and the kept rank with a threshold (1e-4) that
lies well above the FP32 noise floor."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def tn():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tncontract_amd
    tncontract_amd.native_library()
    return tncontract_amd


@pytest.mark.parametrize("njobs,lo,hi", [(6, 16, 40), (40, 16, 160)])
def test_batched_truncate_matches_numpy(tn, njobs, lo, hi):
    from tncontract_amd.batched import batched_truncate, config3_jobs
    jobs = config3_jobs(njobs, seed=njobs, chi_lo=lo, chi_hi=hi)
    Ls = [torch.from_numpy(j[0]).cuda() for j in jobs]
    Rs = [torch.from_numpy(j[1]).cuda() for j in jobs]
    thr, chi = 1e-4, 64
    U, S, Vt, kept, sweeps = batched_truncate(Ls, Rs, rel_threshold=thr, chi=chi)
    for j, (Lm, Rm, sig) in enumerate(jobs):
        C = Lm.astype(np.float64) @ Rm.astype(np.float64)
        s_ref = np.linalg.svd(C, compute_uv=False)
        s = S[j].cpu().numpy()
        big = s_ref > 1e-5 * s_ref[0]
        np.testing.assert_allclose(s[:big.sum()], s_ref[big], rtol=1e-4, atol=1e-6 * s_ref[0])
        k_ref = min(int(np.count_nonzero(s_ref / s_ref[0] > thr)), chi)
        assert kept[j] == k_ref, (j, kept[j], k_ref)
        u, vt = U[j].cpu().numpy().astype(np.float64), Vt[j].cpu().numpy().astype(np.float64)
        k = kept[j]
        rec = (u[:, :k] * s[:k]) @ vt[:k]
        ref = (np.linalg.svd(C)[0][:, :k] * s_ref[:k]) @ np.linalg.svd(C)[2][:k]
        assert np.linalg.norm(rec - ref) <= 1e-4 * np.linalg.norm(C)
        np.testing.assert_allclose(u[:, :k].T @ u[:, :k], np.eye(k), atol=1e-5)
    assert sweeps.max() < 40
