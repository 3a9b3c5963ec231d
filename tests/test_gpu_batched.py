"""Batched uneven contraction + truncation (grouped GEMM + batched Jacobi)
against NumPy FP64 (the oracle's tensordot + LAPACK SVD) on the same FP32
factors.  The product is accumulated in FP64, so the kept rank at the
config's relative threshold 1e-8 equals LAPACK's and the kept singular
values (FP64 outputs) agree to 1e-12 sigma_0; U and V^T are returned in the
factors' FP32, so the truncated factors reconstruct LAPACK's truncation and
are orthonormal to FP32 rounding (1e-6)."""
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


def check_job(Lm, Rm, U, S, Vt, kept, thr, chi):
    C = Lm.astype(np.float64) @ Rm.astype(np.float64)
    u_ref, s_ref, vt_ref = np.linalg.svd(C, full_matrices=False)
    k_ref = int(np.count_nonzero(s_ref / s_ref[0] > thr))
    if chi:
        k_ref = min(k_ref, chi)
    assert kept == k_ref, (kept, k_ref)
    s = S.cpu().numpy()
    np.testing.assert_allclose(s[:kept], s_ref[:kept], rtol=0, atol=1e-12 * s_ref[0])
    u = U.cpu().numpy().astype(np.float64)[:, :kept]
    vt = Vt.cpu().numpy().astype(np.float64)[:kept]
    rec = (u * s[:kept]) @ vt
    ref = (u_ref[:, :kept] * s_ref[:kept]) @ vt_ref[:kept]
    assert np.linalg.norm(rec - ref) <= 1e-6 * np.linalg.norm(C)
    np.testing.assert_allclose(u.T @ u, np.eye(kept), atol=1e-6)
    np.testing.assert_allclose(vt @ vt.T, np.eye(kept), atol=1e-6)


@pytest.mark.parametrize("njobs,lo,hi,chi", [(6, 16, 40, 0), (40, 16, 160, 64), (40, 16, 160, 0)])
def test_batched_truncate_matches_numpy(tn, njobs, lo, hi, chi):
    from tncontract_amd.batched import batched_truncate, config3_jobs
    jobs = config3_jobs(njobs, seed=njobs + chi, chi_lo=lo, chi_hi=hi)
    Ls = [torch.from_numpy(j[0]).cuda() for j in jobs]
    Rs = [torch.from_numpy(j[1]).cuda() for j in jobs]
    U, S, Vt, kept, sweeps = batched_truncate(Ls, Rs, rel_threshold=1e-8, chi=chi)
    for j, (Lm, Rm, sig) in enumerate(jobs):
        check_job(Lm, Rm, U[j], S[j], Vt[j], int(kept[j]), 1e-8, chi)
    assert sweeps.max() < 60


def test_config3_full_size(tn):
    """Config 3 at its stated size: 1024 jobs, chi_l, chi_m, chi_r ~
    U{16..512}, d=2, sigma_k = exp(-k/tau), tau ~ U[4, 64], FP32 factors,
    keep sigma/sigma_0 > 1e-8.  Every job's kept rank is compared with the
    closed form min(rank, #{exp(-k/tau) > 1e-8}) up to the FP32 rounding of
    the factors; 48 jobs, spread over the sizes, against LAPACK in full."""
    from tncontract_amd.batched import batched_truncate, config3_jobs
    jobs = config3_jobs(1024, seed=0)
    Ls = [torch.from_numpy(j[0]).cuda() for j in jobs]
    Rs = [torch.from_numpy(j[1]).cuda() for j in jobs]
    U, S, Vt, kept, sweeps = batched_truncate(Ls, Rs, rel_threshold=1e-8)
    assert sweeps.max() < 60
    for j, (Lm, Rm, sig) in enumerate(jobs):
        # rounding the factors to FP32 moves each sigma by ~1e-7 sigma_0
        lo = int(np.count_nonzero(sig > 1e-8 + 1e-6))
        hi = int(np.count_nonzero(sig > 1e-8 - 1e-6))
        assert lo <= kept[j] <= hi, (j, kept[j], lo, hi)
    order = np.argsort([j[0].shape[0] * j[1].shape[1] for j in jobs])
    for j in order[::len(order) // 48][:48]:
        Lm, Rm, _ = jobs[j]
        check_job(Lm, Rm, U[j], S[j], Vt[j], int(kept[j]), 1e-8, 0)


def test_batched_svd_qr_preconditioned_edges(tn):
    """Jobs at the edges of the Householder QR preconditioning of the batched
    Jacobi (k = min(m, n) > 32 and max(m, n) <= 1024 are preconditioned): the
    smallest preconditioned k (33), exactly 1024 rows, a square 1024 x 1024
    job, a 1025-row job beside them (left unpreconditioned), a rank-10
    product and an all-zero matrix, in one batch against LAPACK.  Exactly
    zero singular values get zero vectors (the library's convention), so the
    orthonormality is checked on the nonzero part."""
    from tncontract_amd.batched import batched_svd
    rng = np.random.default_rng(33)
    shapes = [(33, 40), (40, 33), (1024, 200), (200, 1024), (1024, 1024), (1025, 64), (64, 1025)]
    mats = [rng.standard_normal(s) for s in shapes]
    mats.append(rng.standard_normal((300, 10)) @ rng.standard_normal((10, 100)))
    mats.append(np.zeros((64, 80)))
    U, S, Vt, sweeps = batched_svd([torch.from_numpy(a).cuda() for a in mats])
    for a, u, s, vt in zip(mats, U, S, Vt):
        u, s, vt = u.cpu().numpy(), s.cpu().numpy(), vt.cpu().numpy()
        ref = np.linalg.svd(a, compute_uv=False)
        scale = max(ref[0], 1.0)
        np.testing.assert_allclose(s, ref, rtol=1e-11, atol=1e-12 * scale)
        np.testing.assert_allclose((u * s) @ vt, a, atol=1e-11 * scale)
        r = int(np.count_nonzero(ref > 1e-10 * scale))
        np.testing.assert_allclose(u[:, :r].T @ u[:, :r], np.eye(r), atol=1e-11)
        np.testing.assert_allclose(vt[:r] @ vt[:r].T, np.eye(r), atol=1e-11)
    assert sweeps.max() <= 20, sweeps
