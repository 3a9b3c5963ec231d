"""Learned-rank truncation on the GPU: the MLP kernels' gradients against the
NumPy oracle, training reduces the loss, the certified randomized SVD and its
fallback, and the learned-rank compression sweep (config 4): a rank model
trained on the GPU on the spectra of the method's own sweeps, compared with
the same sweep without it (the full Jacobi SVD at every bond) and with the
oracle."""
import time

import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def tn():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tncontract_amd
    tncontract_amd.native_library()
    return tncontract_amd


def test_mlp_gradient_matches_oracle(tn):
    from tncontract_amd.learned import RankMLP
    rng = np.random.default_rng(0)
    p = mo.init_params(5)
    X = rng.standard_normal((33, 34))
    T = rng.standard_normal(33)
    model = RankMLP(init=p)
    g, y = model.gradient(torch.from_numpy(X.astype(np.float32)).cuda(),
                          torch.from_numpy(T.astype(np.float32)).cuda())
    y_ref, _ = mo.forward(p.astype(np.float32).astype(np.float64), X.astype(np.float32).astype(np.float64))
    np.testing.assert_allclose(y.cpu().numpy(), y_ref, rtol=1e-4, atol=1e-4)
    g_ref = mo.grad(p.astype(np.float32).astype(np.float64), X.astype(np.float32).astype(np.float64), T)
    gg = g.cpu().numpy()
    assert np.linalg.norm(gg - g_ref) <= 1e-4 * np.linalg.norm(g_ref)


def test_mlp_training_reduces_loss(tn):
    from tncontract_amd.learned import RankMLP
    rng = np.random.default_rng(1)
    X = rng.uniform(0, 1, (512, 34)).astype(np.float32)
    T = (0.3 * X[:, 0] + 0.5 * X[:, 5] * X[:, 6]).astype(np.float32)
    model = RankMLP(seed=2)
    losses = model.fit(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda(), epochs=120, lr=3e-3)
    assert losses[-1] < 0.2 * losses[0], losses


def lowrank(m, n, r, tau, seed):
    rng = np.random.default_rng(seed)
    u, _ = np.linalg.qr(rng.standard_normal((m, r)))
    v, _ = np.linalg.qr(rng.standard_normal((n, r)))
    s = np.exp(-np.arange(r) / tau)
    return (u * s) @ v.T, s


def test_certified_svd_sketch_and_fallback(tn):
    from tncontract_amd.learned import truncated_svd_learned
    a, s = lowrank(600, 300, 40, 3.0, 0)          # exact rank 40, sigma_39 ~ 2e-6
    A = torch.from_numpy(a).cuda()
    U, S, Vt, kept, info = truncated_svd_learned(A, 40, rel_threshold=1e-12)
    assert info["path"] == "sketch" and info["residual"] <= 1e-12
    assert kept == 40
    np.testing.assert_allclose(S.cpu().numpy(), s, rtol=1e-9, atol=1e-14)
    # an under-predicted rank saturates the sketch: full SVD, same result
    U2, S2, Vt2, kept2, info2 = truncated_svd_learned(A, 5, rel_threshold=1e-12)
    assert info2["path"] == "fallback" and kept2 == 40
    np.testing.assert_allclose(S2.cpu().numpy(), s, rtol=1e-9, atol=1e-14)
    # a sketch as large as half the matrix is not tried
    _, _, _, _, info3 = truncated_svd_learned(A, 200, rel_threshold=1e-12)
    assert info3["path"] == "full"


def _to_oracle(g):
    from oracle import tn_oracle as o
    return o.OMPS([o.OTensor(t.numpy().astype(np.float64), t.labels) for t in g],
                  g.left_label, g.right_label, g.phys_label)


def _rel_dist2(a, b):
    from oracle import tn_oracle as o
    nb = o.inner_product_mps(b, b)
    return float((o.inner_product_mps(a, a) + nb - 2 * o.inner_product_mps(a, b)) / nb)


def test_learned_rank_sweep_config4(tn):
    """Config 4: bond matrices up to 2048 x 1024 (d=4, input bond 1024,
    compressed to chi=512 at sigma/sigma_0 > 1e-8; FP32 data, FP64
    arithmetic).  The rank model is trained on the GPU on the spectra of
    full-SVD sweeps of other states (the same family with other decay rates,
    and config 1's flat spectra); each bond's previous-sweep spectrum comes
    from a sibling state with the same decay rates.  The learned sweep must
    give the oracle's bond dimensions and state (<= 1e-10), and factor fewer
    columns than the full SVDs of the sweep without the network."""
    import time
    from oracle import tn_oracle as o
    from tncontract_amd import learned as lr
    chi, thr = 512, 1e-8
    samples = []
    for seed in (1, 2, 3, 4):
        prev, cur, _ = lr.decaying_pair(seed)
        samples.append(lr.sweep_samples(prev, cur, chi, thr)[:2])
    for seed in (1, 2):
        prev, cur = lr.flat_pair(seed, 20, 2, 32)
        samples.append(lr.sweep_samples(prev, cur, 16, 1e-10)[:2])
    model, losses = lr.train_rank_model(samples)
    assert losses[-1] < 0.5 * losses[0], losses
    prev, cur, taus = lr.decaying_pair(99)
    X, T, spectra_prev, needed_cur = lr.sweep_samples(prev, cur, chi, thr)
    pred = model.predict_rank(X, chi)
    rank_err = float(np.mean(np.abs(pred - T * chi)) / chi)
    # the sweep without the network: full Jacobi SVD at every bond
    state = lr._state(cur)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref_sweep, _, kept_full = lr.record_sweep(state, chi, thr)
    torch.cuda.synchronize()
    t_full = time.perf_counter() - t0
    st = lr.LearnedSweepStats()
    t0 = time.perf_counter()
    out = lr.learned_svd_compress(state, chi, thr, model, spectra_prev, stats=st)
    torch.cuda.synchronize()
    t_learned = time.perf_counter() - t0
    print("config 4 learned sweep: %.3fs vs full-SVD sweep %.3fs; paths %s; rank error %.3f; "
          "columns factored %d of %d" % (t_learned, t_full, st.path, rank_err, st.work_cols,
                                         st.full_cols))
    assert out.bonddims() == ref_sweep.bonddims()
    r = o.OMPS([o.OTensor(a.astype(np.float64), ["phys", "left", "right"]) for a in cur])
    r.svd_compress(chi=chi, threshold=thr)
    assert out.bonddims() == r.bonddims()
    assert _rel_dist2(_to_oracle(out), r) < 1e-10
    assert _rel_dist2(_to_oracle(ref_sweep), r) < 1e-10
    assert rank_err < 0.2
    assert any(p == "sketch" for p in st.path.values())
    assert st.work_cols < st.full_cols
