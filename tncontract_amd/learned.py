"""Learned-rank truncation: a small MLP predicts, from the previous sweep's
spectrum at each bond, the rank a sketch of that bond needs to certify the
keep rule, so a randomized range-finder SVD
(Halko, Martinsson & Tropp 2011: Gaussian sketch of r + p columns, optional
power iterations with Householder re-orthogonalisation; the sweep uses
p = 50 and none) can replace the full Jacobi SVD in the truncating sweep of
svd_compress.  The sketch is accepted
only when it certifies the reference's keep rule (sigma / sigma_0 >
threshold): an upper bound on the spectral norm of the residual
A - U S V^T, holding with probability at least 1 - 1e-10 (the posterior
estimate from 20 Gaussian test vectors, Halko, Martinsson & Tropp, Lemma
4.1), must not exceed threshold * sigma_0 (so every singular value the sketch missed lies below
the threshold), and no sketched value may lie within that bound below the
threshold (so no kept value can hide under the estimate).  Otherwise the full
one-sided Jacobi SVD runs (the fallback), exactly as in svd_compress.

Network: 34 -> 64 -> 64 -> 1 (ReLU, FP32), trained on the GPU by the
library's own forward / backward / Adam kernels (tn_mlp_*).  Input features:
the 32 leading sigma/sigma_0 of the previous sweep at the bond (zero padded),
log(chi) and the bond position i/N.  Output: the rank a certifying sketch
needs (the count of sigma / sigma_0 above threshold / CERT_FACTOR) / chi.

The reference has no counterpart: it always runs the full SVD
(tensor.py:915, onedim_core.py:317).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from tncontract_amd import _lib as L
from tncontract_amd import tensor as tsr

NFEAT = 34


def _bind():
    lib = L.lib()
    if not hasattr(lib, "_tn_mlp_bound"):
        fp, vp, i32, flt = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        lib.tn_mlp_param_count.restype = i32
        lib.tn_mlp_forward.restype = i32
        lib.tn_mlp_forward.argtypes = [fp, i32, fp, fp, fp, fp, vp]
        lib.tn_mlp_backward.restype = i32
        lib.tn_mlp_backward.argtypes = [fp, fp, i32, fp, fp, fp, fp, fp, fp, vp]
        lib.tn_adam_step.restype = i32
        lib.tn_adam_step.argtypes = [fp, fp, fp, fp, i32, flt, flt, flt, flt, i32, vp]
        lib.tn_randn.restype = i32
        lib.tn_randn.argtypes = [vp, i32, ctypes.c_int64, ctypes.c_ulonglong, vp]
        lib._tn_mlp_bound = True
    return lib


def features(rel_spectrum, chi, position):
    """34 features of one bond: 32 leading sigma/sigma_0 (zero padded),
    log(chi), position in [0, 1]."""
    f = np.zeros(NFEAT, dtype=np.float32)
    s = np.asarray(rel_spectrum, dtype=np.float64)[:32]
    f[:len(s)] = s
    f[32] = math.log(max(chi, 1))
    f[33] = position
    return f


class RankMLP:
    """The rank-prediction network; parameters live on the GPU."""

    def __init__(self, seed=0, init=None):
        lib = _bind()
        self.np = lib.tn_mlp_param_count()
        if init is None:
            from tncontract_amd._mlp_init import he_init  # host-side initial weights
            init = he_init(seed)
        self.P = torch.from_numpy(np.asarray(init, dtype=np.float32)).to(L.device())
        self.m = torch.from_numpy(np.zeros(self.np, dtype=np.float32)).to(L.device())
        self.v = torch.from_numpy(np.zeros(self.np, dtype=np.float32)).to(L.device())
        self.step = 0

    def forward(self, X, keep_activations=False):
        lib = _bind()
        X = L.contiguous(X)
        B = X.shape[0]
        Y = torch.empty((B,), dtype=torch.float32, device=X.device)
        H1 = H2 = None
        if keep_activations:
            H1 = torch.empty((B, 64), dtype=torch.float32, device=X.device)
            H2 = torch.empty((B, 64), dtype=torch.float32, device=X.device)
        L.check(lib.tn_mlp_forward(L.ptr(X), B, L.ptr(self.P), L.ptr(Y),
                                   L.ptr(H1) if H1 is not None else None,
                                   L.ptr(H2) if H2 is not None else None, L.stream_ptr()))
        return Y, H1, H2

    def gradient(self, X, T):
        """d/dP of mean((y - t)^2) on the GPU (returns the device gradient)."""
        lib = _bind()
        Y, H1, H2 = self.forward(X, keep_activations=True)
        B = X.shape[0]
        grad = torch.empty((self.np,), dtype=torch.float32, device=X.device)
        scratch = torch.empty((B, self.np), dtype=torch.float32, device=X.device)
        L.check(lib.tn_mlp_backward(L.ptr(X), L.ptr(T), B, L.ptr(self.P), L.ptr(Y), L.ptr(H1),
                                    L.ptr(H2), L.ptr(grad), L.ptr(scratch), L.stream_ptr()))
        return grad, Y

    def fit(self, X, T, epochs=300, lr=3e-3, batch=256, seed=0):
        """Adam on the GPU.  X: (n, 34) float32 device tensor, T: (n,)."""
        lib = _bind()
        n = X.shape[0]
        order = np.random.default_rng(seed)
        losses = []
        for ep in range(epochs):
            perm = order.permutation(n)
            for b0 in range(0, n, batch):
                idx = torch.from_numpy(perm[b0:b0 + batch]).to(X.device)
                xb = L.contiguous(X.index_select(0, idx))
                tb = L.contiguous(T.index_select(0, idx))
                grad, Y = self.gradient(xb, tb)
                self.step += 1
                L.check(lib.tn_adam_step(L.ptr(self.P), L.ptr(grad), L.ptr(self.m), L.ptr(self.v),
                                         self.np, lr, 0.9, 0.999, 1e-8, self.step, L.stream_ptr()))
            if ep == epochs - 1 or ep % 50 == 0:
                Y, _, _ = self.forward(X)
                d = (Y - T).cpu().numpy()
                losses.append(float(np.mean(d * d)))
        return losses

    def predict_rank(self, feats, chi):
        X = torch.from_numpy(np.asarray(feats, dtype=np.float32).reshape(-1, NFEAT)).to(L.device())
        Y, _, _ = self.forward(X)
        r = np.rint(Y.cpu().numpy() * chi).astype(int)
        return np.clip(r, 1, 4 * chi)


# ---------------------------------------------------------------- randomized SVD
def randn(shape, dtype=torch.float64, seed=0):
    lib = _bind()
    out = torch.empty(tuple(shape), dtype=dtype, device=L.device())
    L.check(lib.tn_randn(L.ptr(out), L.dtype_code(out), out.numel(), int(seed) & (2**64 - 1),
                         L.stream_ptr()))
    return out


def _orth(Y):
    """Orthonormal basis of the columns of the tall device matrix Y by
    Householder QR (the power iterations leave Y too ill-conditioned for
    Cholesky QR: measured, trying it first only added time)."""
    q, _ = tsr._qr_device(L.convert(Y, torch.float64), 1)
    return q


def rsvd(A, rank, oversample=10, power_iters=2, seed=0):
    """Randomized thin SVD of the m x n device matrix A (FP64 arithmetic; FP32
    data is exact in FP64) with p = rank + oversample sketch columns and
    `power_iters` power iterations, each application of A or A^T followed by a
    re-orthonormalisation (_orth).  Returns (U m x p, s p, Vt p x n)."""
    m, n = A.shape
    p = min(rank + oversample, m, n)
    A64 = L.convert(A, torch.float64)
    con = tsr._contract_data
    Om = randn((n, p), dtype=torch.float64, seed=seed)
    Q = _orth(con(A64, [1], Om, [0]))          # m x p
    for _ in range(power_iters):
        Z = _orth(con(A64, [0], Q, [0]))       # n x p  (A^T Q)
        Q = _orth(con(A64, [1], Z, [0]))       # m x p
    Bm = con(Q, [0], A64, [0])                 # p x n = Q^T A
    ub, s, vt = tsr._svd_device(Bm, 1)
    U = con(Q, [1], ub, [0])                   # m x p
    return U, s, vt


# posterior bound on the sketch's residual (Halko, Martinsson and Tropp,
# SIAM Rev. 53 (2011), Lemma 4.1): for Gaussian w_1..w_r and alpha > 1,
# ||(I - Q Q^T) A||_2 <= alpha sqrt(2/pi) max_i ||(I - Q Q^T) A w_i|| with
# probability at least 1 - alpha^-r; r = 20 test vectors with
# alpha = 10^(1/2) keep the failure probability at 1e-10 with a bound 3.2
# times tighter than r = 10, alpha = 10
POSTERIOR_TESTS = 20
POSTERIOR_FACTOR = 10.0 ** 0.5 * np.sqrt(2.0 / np.pi)


def residual_bound(A, U, ntest=POSTERIOR_TESTS, seed=1):
    """Upper bound, with probability at least 1 - 10^-(ntest/2), on the spectral
    norm of the sketch's residual A - U U^T A (U: the sketch's orthonormal
    m x p basis; A - U diag(s) Vt equals it, s and Vt being the full SVD of
    U^T A), from ntest Gaussian test vectors: three thin products, no
    factorisation."""
    con = tsr._contract_data
    A64 = L.convert(A, torch.float64)
    W = randn((A64.shape[1], ntest), dtype=torch.float64, seed=seed)
    Y = con(A64, [1], W, [0])                                   # m x ntest
    E = L.axpby(Y, con(U, [1], con(U, [0], Y, [0]), [0]), 1.0, -1.0)
    col2 = torch.diagonal(con(E, [0], E, [0])).cpu().numpy()   # ||E_i||^2
    return POSTERIOR_FACTOR * float(np.sqrt(max(float(np.max(col2)), 0.0)))


def certified_svd(A, rank, threshold, oversample=10, power_iters=2, seed=0, full_fraction=0.5):
    """Truncation SVD of the m x n device matrix A for the keep rule
    sigma / sigma_0 > threshold, by the randomized SVD at the predicted
    `rank` when it certifies that rule, else by the full Jacobi SVD.
    Returns (U, s, Vt, info): U m x p, s p (descending, FP64), Vt p x n with
    p = the sketch size or min(m, n); every singular value of A above
    threshold * sigma_0 is among s.  info: {"path": "sketch" | "fallback" |
    "full", "rank": rank, "residual": the residual bound / sigma_0 or None}.
    "full": the sketch would hold more than full_fraction * min(m, n) columns,
    so the full SVD runs without trying it."""
    m, n = A.shape
    kmin = min(m, n)
    info = {"path": "full", "rank": int(rank), "residual": None}
    if rank + oversample <= full_fraction * kmin:
        U, s, vt = rsvd(A, int(rank), oversample, power_iters, seed)
        sh = s.cpu().numpy()
        s0 = sh[0] if sh.size else 0.0
        if s0 > 0.0 and np.all(np.isfinite(sh)):
            bound = residual_bound(A, U, seed=seed + 1)
            info["residual"] = bound / s0
            cut = threshold * s0
            certified = (bound <= cut and int(np.count_nonzero(sh > cut)) < sh.size
                         and not np.any((sh <= cut) & (sh > cut - bound)))
            if certified:
                info["path"] = "sketch"
                return U, s, vt, info
        info["path"] = "fallback"
    u, s, vt = tsr._svd_device(L.convert(A, torch.float64) if A.dtype != torch.float64 else A, 1)
    return u, s, vt, info


def truncated_svd_learned(A, predicted_rank, rel_threshold=1e-8, chi=0, **kw):
    """certified_svd plus the keep rule: returns (U, S, Vt, kept, info) with
    the first `kept` triplets (sigma / sigma_0 > rel_threshold, at most chi)."""
    U, s, vt, info = certified_svd(A, predicted_rank, rel_threshold, **kw)
    sh = s.cpu().numpy()
    keep = int(np.count_nonzero(sh / sh[0] > rel_threshold)) if sh.size and sh[0] > 0 else 0
    if chi:
        keep = min(keep, chi)
    return U[:, :keep], s[:keep], vt[:keep], keep, info


# ---------------------------------------------------------------- learned sweep
def _bond_of(pos, N):
    """Bond index (bond b joins sites b - 1 and b) of the SVD at position pos
    of the reversed chain in svd_compress's right sweep."""
    return N - 1 - pos


class LearnedSweepStats:
    """Per-bond record of a learned-rank compression sweep."""

    def __init__(self):
        self.predicted = {}   # bond -> predicted rank (before the margin)
        self.kept = {}        # bond -> kept rank
        self.path = {}        # bond -> "sketch" | "fallback" | "full"
        self.spectra = {}     # bond -> sigma / sigma_0 the truncation saw
        self.work_cols = 0    # columns factored: sketches tried plus full SVDs run
        self.full_cols = 0    # min(m, n) of every bond matrix (the sweep without the model)

    @property
    def fallbacks(self):
        return sum(1 for p in self.path.values() if p == "fallback")


# The sketch shape of the learned sweep: no power iteration and 50
# oversampling columns.  Measured on config 4 (against the same sweep with
# the full SVD at every bond): q = 2, p = 10: 0.87x, 2 fallbacks; q = 1,
# p = 20: 1.01x, 1; q = 0, p = 30: 1.12x, 1; q = 0, p = 50: 1.20x, none;
# p = 70 or 100 send more bonds to the full SVD (the sketch would exceed
# half of min(m, n)): 1.12x, 1.06x.  The decaying bond spectra make the
# oversampling, not the power iterations (two Householder QRs each), the
# cheaper way to certify.  The predicted rank's widening (scale 1.2,
# margin 16) and the largest sketch tried (half of min(m, n)) measured best
# with this shape: margin 0 / scale 1.0 gave 0.95x (4 fallbacks), margin 8 /
# scale 1.1 1.12x (1), sketches up to 0.7 min(m, n) 1.18x.
SWEEP_OVERSAMPLE = 50
SWEEP_POWER_ITERS = 0


def learned_svd_compress(mps, chi, threshold, model, prev_spectra, stats=None, margin=16,
                         scale=1.2, oversample=SWEEP_OVERSAMPLE, power_iters=SWEEP_POWER_ITERS,
                         seed=0):
    """svd_compress(chi, threshold) (onedim_core.py:463-484) of a copy of
    `mps` with each truncating SVD of the right sweep done by certified_svd
    at the rank `model` predicts from prev_spectra[bond] (sigma / sigma_0 of
    the previous sweep at that bond; bonds without one use chi), widened to
    scale * rank + margin: an under-predicted rank costs a sketch and the
    full SVD, an over-predicted one only a wider sketch.  The keep
    rule, the step order and the outputs are svd_compress's; a bond whose
    sketch is not certified falls back to the full Jacobi SVD.  Returns the
    compressed MPS."""
    N = len(mps)
    out = mps.copy()
    bonds = list(range(1, N))
    feats = [features(prev_spectra.get(b, np.ones(32)), chi, b / N) for b in bonds]
    ranks = dict(zip(bonds, (int(r) for r in model.predict_rank(feats, chi))))
    for b in bonds:
        if b not in prev_spectra:
            ranks[b] = chi

    def svd_hook(site, row_labels, svd_label, pos):
        b = _bond_of(pos, N)
        t, nrow = tsr._row_view(site, row_labels, move_all=False)
        shape = t.shape
        m = int(np.prod(shape[:nrow]))
        A = L.materialize(t.data).view(m, -1)
        r = int(np.ceil(scale * ranks.get(b, chi))) + margin
        U, s, vt, info = certified_svd(A, r, threshold, oversample, power_iters, seed + b)
        k = s.numel()
        if stats is not None:
            stats.predicted[b] = ranks.get(b, chi)
            stats.path[b] = info["path"]
            stats.full_cols += min(A.shape)
            if info["path"] != "full":
                stats.work_cols += min(r + oversample, min(A.shape))
            if info["path"] != "sketch":
                stats.work_cols += min(A.shape)
            sh = s.cpu().numpy()
            rel = sh / sh[0] if sh.size and sh[0] > 0 else sh
            stats.spectra[b] = rel
            kept = int(np.count_nonzero(rel > threshold))
            stats.kept[b] = min(kept, chi) if chi else kept
        U, vt = L.convert(U, site.dtype), L.convert(vt, site.dtype)
        Ut = tsr.Tensor._wrap(U.view(tuple(shape[:nrow]) + (k,)),
                              list(row_labels) + [svd_label + "in"])
        col_labels = [x for x in t.labels if x not in row_labels]
        Vt_ = tsr.Tensor._wrap(vt.view((k,) + tuple(shape[nrow:])), [svd_label + "out"] + col_labels)
        St = tsr.Tensor._wrap(L.diag_embed(s, U.dtype), [svd_label + "out", svd_label + "in"])
        return Ut, St, Vt_, s

    out.svd_compress(chi=chi, threshold=threshold, _svd=svd_hook)
    return out


# A sketch certifies the keep rule when its residual bound is below
# threshold * sigma_0, i.e. when it holds every singular
# value down to about threshold / CERT_FACTOR: the rank the model predicts
# (the sketch's size less its oversampling) is that count, not the kept rank.
CERT_FACTOR = 16.0


def record_sweep(mps, chi, threshold):
    """svd_compress(chi, threshold) of a copy of `mps` by the full Jacobi SVD,
    recording each bond's sigma / sigma_0 and the rank a certifying sketch of
    it needs, #{sigma / sigma_0 > threshold / CERT_FACTOR} (the training data
    of the rank model and the "previous sweep" of the next one).  Returns
    (compressed MPS, {bond: rel spectrum}, {bond: sketch rank})."""
    N = len(mps)
    spectra, needed = {}, {}

    def svd_hook(site, row_labels, svd_label, pos):
        U, S, V, s = tsr._tensor_svd(site, row_labels, svd_label)
        b = _bond_of(pos, N)
        sh = s.cpu().numpy()
        rel = sh / sh[0] if sh.size and sh[0] > 0 else sh
        spectra[b] = rel
        needed[b] = int(np.count_nonzero(rel > threshold / CERT_FACTOR))
        return U, S, V, s

    out = mps.copy()
    out.svd_compress(chi=chi, threshold=threshold, _svd=svd_hook)
    return out, spectra, needed


def training_samples(prev_spectra, ranks, chi, N):
    """(features of the previous sweep's spectrum at each bond, rank / chi of
    this sweep) for every bond both sweeps saw."""
    X, T = [], []
    for b, k in ranks.items():
        if b in prev_spectra:
            X.append(features(prev_spectra[b], chi, b / N))
            T.append(k / float(chi))
    return np.array(X, dtype=np.float32).reshape(-1, NFEAT), np.array(T, dtype=np.float32)


# ---------------------------------------------------------------- workload and training
def _state(arrs, dtype=torch.float64):
    from tncontract_amd.onedim import MatrixProductState
    return MatrixProductState([tsr.Tensor(L.convert(torch.from_numpy(np.ascontiguousarray(a))
                                                    .to(L.device()), dtype),
                                          ["phys", "left", "right"]) for a in arrs])


def decaying_pair(seed, N=16, d=4, chi_in=1024, tau_range=(2.0, 16.0)):
    """Two seeded states of the learned-rank workload with the same bond decay
    rates (one tau ~ U[tau_range] per bond) and different isometries: the
    previous sweep's state and the current one.  FP32 data (host arrays).
    The default rates are those of fast-decaying (gapped, ground-state-like)
    spectra, whose truncation at 1e-8 keeps ~18 tau of the 1024 values: the
    regime a rank-sized sketch is for."""
    from tncontract_amd.onedim.onedim_utils import decaying_mps_arrays
    rng = np.random.default_rng(seed)
    taus = rng.uniform(tau_range[0], tau_range[1], size=N - 1)
    prev = decaying_mps_arrays(N, d, chi_in, taus, seed=2 * seed + 1)
    cur = decaying_mps_arrays(N, d, chi_in, taus, seed=2 * seed + 2)
    return prev, cur, taus


def flat_pair(seed, N, d, chi_in):
    """Two seeded random states (i.i.d. N(0,1) sites, normalised) of the same
    shape: the flat spectra of the plain configurations."""
    from tncontract_amd.onedim.onedim_utils import random_mps_arrays
    return random_mps_arrays(N, d, chi_in, seed=2 * seed + 1), random_mps_arrays(N, d, chi_in,
                                                                                 seed=2 * seed + 2)


def sweep_samples(prev_arrs, cur_arrs, chi, threshold):
    """Training samples from two full-SVD compression sweeps: the previous
    sweep's spectra as features, the current sweep's certifying sketch ranks
    as targets.  Returns (X, T, prev spectra, current sketch ranks)."""
    N = len(cur_arrs)
    _, spectra_prev, _ = record_sweep(_state(prev_arrs), chi, threshold)
    _, _, needed_cur = record_sweep(_state(cur_arrs), chi, threshold)
    X, T = training_samples(spectra_prev, needed_cur, chi, N)
    return X, T, spectra_prev, needed_cur


def train_rank_model(samples, epochs=400, lr=3e-3, seed=0):
    """RankMLP fitted on the GPU to (X, T) sample sets from sweeps.  Returns
    (model, losses)."""
    X = np.concatenate([s[0] for s in samples])
    T = np.concatenate([s[1] for s in samples])
    model = RankMLP(seed=seed)
    losses = model.fit(torch.from_numpy(X).to(L.device()), torch.from_numpy(T).to(L.device()),
                       epochs=epochs, lr=lr, batch=64, seed=seed)
    return model, losses
