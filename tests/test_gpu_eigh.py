"""PSD eigensolver (tn_eigh_psd) against LAPACK's eigh (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def T():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from tncontract_amd import tensor
    return tensor


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def check_eig(K, u, lam, nkeep, rel=1e-11):
    """Eigenpairs 0..nkeep-1: residual, orthonormality, eigenvalues vs eigh."""
    ref = np.linalg.eigvalsh(K)[::-1]
    scale = max(abs(ref[0]), 1e-300)
    assert np.all(np.diff(lam) <= 1e-12 * scale)                      # descending
    assert np.abs(lam - ref).max() <= rel * scale * max(1.0, np.sqrt(len(lam)) / 8)
    uk = u[:, :nkeep]
    res = K @ uk - uk * lam[:nkeep]
    assert np.linalg.norm(res) <= rel * scale * np.sqrt(K.shape[0])
    assert np.abs(uk.T @ uk - np.eye(nkeep)).max() <= 1e-11


@pytest.mark.parametrize("n", [1, 2, 5, 17, 64, 65, 100, 257, 600])
def test_eigh_psd_full_rank(T, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n + 3, n))
    K = A.T @ A
    u, lam = T._eigh_psd_device(dev(K), 1)
    check_eig(K, u.cpu().numpy(), lam.cpu().numpy(), n)


@pytest.mark.parametrize("n,r", [(64, 20), (300, 150), (513, 100)])
def test_eigh_psd_rank_deficient(T, n, r):
    rng = np.random.default_rng(n + r)
    A = rng.standard_normal((r, n))
    K = A.T @ A                                      # rank r, n - r zero eigenvalues
    u, lam = T._eigh_psd_device(dev(K), 1)
    lam = lam.cpu().numpy()
    check_eig(K, u.cpu().numpy(), lam, r)
    assert np.abs(lam[r:]).max() <= 1e-12 * lam[0] * np.sqrt(n)


def test_eigh_psd_graded_spectrum(T):
    # eigenvalues 1 .. 1e-16: the top of the spectrum to relative accuracy
    n = 200
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = 10.0 ** -np.linspace(0, 16, n)
    K = (Q * ev) @ Q.T
    K = 0.5 * (K + K.T)
    u, lam = T._eigh_psd_device(dev(K), 1)
    u, lam = u.cpu().numpy(), lam.cpu().numpy()
    big = int(np.count_nonzero(ev > 1e-8))
    assert np.abs(lam[:big] - ev[:big]).max() <= 1e-12
    # top subspace: K's own rounding (~1e-16 absolute) moves it by ~1e-16 / gap
    top = int(np.count_nonzero(ev > 1e-4))
    P = u[:, :top] @ u[:, :top].T
    Pref = Q[:, :top] @ Q[:, :top].T
    assert np.linalg.norm(P - Pref) <= 1e-7


def test_eigh_psd_strided_and_fp32(T):
    rng = np.random.default_rng(9)
    A = rng.standard_normal((90, 80))
    K = A.T @ A
    Kt = dev(K).t()                                  # strided view (K is symmetric)
    u, lam = T._eigh_psd_device(Kt, 1)
    check_eig(K, u.cpu().numpy(), lam.cpu().numpy(), 80)
    K32 = K.astype(np.float32)
    u, lam = T._eigh_psd_device(dev(K32), 1)
    check_eig(K32.astype(np.float64), u.cpu().numpy(), lam.cpu().numpy(), 80, rel=1e-10)


def test_eigh_psd_zero_matrix(T):
    K = np.zeros((40, 40))
    u, lam = T._eigh_psd_device(dev(K), 1)
    assert np.all(lam.cpu().numpy() == 0.0)
    assert np.all(np.isfinite(u.cpu().numpy()))


def test_eigh_psd_four_index_view(T):
    # the structured sweep's layout: K[k', t', k, t] read as rows (t', k'), cols (t, k)
    rng = np.random.default_rng(4)
    r, d = 30, 2
    C = rng.standard_normal((50, r * d))
    K2 = C.T @ C                                     # rows/cols (t, k) fused as t*r + k
    K4 = K2.reshape(d, r, d, r).transpose(1, 0, 3, 2).copy()   # (k', t', k, t)
    view = dev(K4).permute(1, 0, 3, 2)
    u, lam = T._eigh_psd_device(view, 2)
    check_eig(K2, u.cpu().numpy(), lam.cpu().numpy(), r * d)


def test_eigh_psd_fp32_rounded_rank_deficient(T):
    # K formed in FP32 from a rank-r factor is indefinite at ~eps_32 |K|; passed
    # as FP32, the shift covers that rounding and the top r pairs come out
    rng = np.random.default_rng(12)
    n, r = 96, 40
    A = rng.standard_normal((r, n)).astype(np.float32)
    K32 = A.T @ A
    K = 0.5 * (K32.astype(np.float64) + K32.astype(np.float64).T)
    assert np.linalg.eigvalsh(K)[0] < 0.0                     # indefinite as given
    u, lam = T._eigh_psd_device(dev(K32), 1)
    u, lam = u.cpu().numpy(), lam.cpu().numpy()
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(lam))
    ev, V = np.linalg.eigh(K)
    ev, V = ev[::-1], V[:, ::-1]
    assert np.abs(lam[:r] - ev[:r]).max() <= 1e-6 * ev[0]
    assert np.abs(lam[r:]).max() <= 1e-4 * ev[0]
    P = u[:, :r] @ u[:, :r].T
    assert np.linalg.norm(P - V[:, :r] @ V[:, :r].T) <= 1e-5
