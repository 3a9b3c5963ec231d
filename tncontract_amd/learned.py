"""Learned-rank truncation: a small MLP predicts each bond's kept rank from
the previous sweep's spectrum at that bond, so a randomized range-finder SVD
(Halko, Martinsson & Tropp 2011: Gaussian sketch, r + 10 columns, 2 power
iterations with Householder re-orthogonalisation) can replace the full
Jacobi SVD.  The full SVD is the fallback whenever the a-posteriori residual
||A - U S V^T||_F / ||A||_F (FP64, explicit) exceeds the tolerance.

Network: 34 -> 64 -> 64 -> 1 (ReLU, FP32), trained on the GPU by the
library's own forward / backward / Adam kernels (tn_mlp_*).  Input features:
the 32 leading sigma/sigma_0 of the previous sweep at the bond (zero padded),
log(chi) and the bond position i/N.  Output: kept rank / chi.

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
        return np.clip(r, 1, chi)


# ---------------------------------------------------------------- randomized SVD
def randn(shape, dtype=torch.float64, seed=0):
    lib = _bind()
    out = torch.empty(tuple(shape), dtype=dtype, device=L.device())
    L.check(lib.tn_randn(L.ptr(out), L.dtype_code(out), out.numel(), int(seed) & (2**64 - 1),
                         L.stream_ptr()))
    return out


def _orth(Y):
    q, _ = tsr._qr_device(L.convert(Y, torch.float64), 1)
    return q


def rsvd(A, rank, oversample=10, power_iters=2, seed=0, precision="native"):
    """Randomized thin SVD of the m x n device matrix A with rank+oversample
    sketch columns.  Returns (U, S, Vt, residual) with the explicit FP64
    residual ||A - U S Vt||_F / ||A||_F."""
    m, n = A.shape
    p = min(rank + oversample, m, n)
    A64 = L.convert(A, torch.float64)          # FP32 data is exact in FP64
    con = lambda x, ax1, y, ax2: tsr._contract_data(x, ax1, y, ax2, precision)  # noqa: E731
    Om = randn((n, p), dtype=torch.float64, seed=seed)
    Y = con(A64, [1], Om, [0])                 # m x p
    for _ in range(power_iters):
        Q = _orth(Y)
        Z = con(A64, [0], Q, [0])              # n x p  (A^T Q)
        Qz = _orth(Z)
        Y = con(A64, [1], Qz, [0])             # m x p
    Q = _orth(Y)                               # m x p, FP64
    Bm = tsr._contract_data(Q, [0], A64, [0])  # p x n = Q^T A (FP64)
    ub, s, vt = tsr._svd_device(Bm, 1)
    U = tsr._contract_data(Q, [1], ub, [0])    # m x p
    # explicit residual: A - U diag(s) Vt
    Us = tsr._contract_data(U, [1], L.diag_embed(s), [0])
    rec = tsr._contract_data(Us, [1], vt, [0])
    res = math.sqrt(L.sumsq(L.axpby(A64, rec, 1.0, -1.0))) / max(math.sqrt(L.sumsq(A64)), 1e-300)
    return U, s, vt, res


def truncated_svd_learned(A, predicted_rank, rel_threshold=1e-8, chi=0, tol=1e-8,
                          oversample=10, power_iters=2, seed=0):
    """Truncated SVD of A using the predicted rank; falls back to the full
    one-sided Jacobi SVD when the randomized residual exceeds `tol` (or the
    sketch would not be smaller than the matrix).  Returns
    (U, S, Vt, kept, info) with info = {"fallback": bool, "residual": float}."""
    m, n = A.shape
    kmin = min(m, n)
    info = {"fallback": False, "residual": None, "rank": int(predicted_rank)}
    if predicted_rank + oversample < kmin:
        U, s, vt, res = rsvd(A, int(predicted_rank), oversample, power_iters, seed)
        info["residual"] = res
        if res > tol:
            info["fallback"] = True
    else:
        info["fallback"] = True
    if info["fallback"]:
        U, s, vt = tsr._svd_device(L.convert(A, torch.float64), 1)
    sh = s.cpu().numpy()
    keep = int(np.count_nonzero(sh / sh[0] > rel_threshold)) if sh.size and sh[0] > 0 else 0
    if chi:
        keep = min(keep, chi)
    return U[:, :keep], s[:keep], vt[:keep], keep, info
