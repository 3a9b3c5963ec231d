"""Structured MPO application + compression in one sweep pair (the fast path).

``apply_mpo_compress(mps, mpo, chi, threshold)`` returns the same state as

    out = contract_mps_mpo(mps, mpo)          # onedim_core.py:1691
    out.svd_compress(chi=chi, threshold=threshold)   # onedim_core.py:463

(same bond dimensions, labels and axis orders; the site tensors agree up to
the gauge of each bond, which the state does not depend on), without ever
forming the chi*D-bond chain or factorising its 2chi*D x chi*D sites.

How: svd_compress = QR left sweep + truncating SVD right sweep.  The QR sweep
only makes the left part of the chain an isometry; the truncation at each
bond then depends on the left part through its Gram matrix alone.  So:

1. Left environments (Gram matrices of the left part of the applied state),
   from the factored site A = B (x) W, never built:
       E_{i+1}[b,beta,b',beta'] = sum E_i[a,alpha,a',alpha'] B[s,a,b] B[s',a',b']
                                  * (sum_t W[alpha,beta,t,s] W[alpha',beta',t,s'])
   as three contractions (2 GEMMs of ~2 chi^3 D^2 d flops each + one K=D^2 d
   contraction), each normalised on the device.
2. Right-to-left truncation carrying P_{i+1} (right bond of the fused site x
   kept rank):  C_i = A_i P_{i+1};  K_i = C_i^T E_i C_i  (d r x d r, FP64);
   K_i = Y diag(sigma^2) Y^T (Cholesky-preconditioned one-sided Jacobi, FP64
   rotations: tn_eigh_psd); keep the
   reference's rule (sigma/sigma_0 > threshold, at most chi, at most the QR
   bond dimension the reference's left sweep would have produced);  site i =
   Y_kept (right-isometric, labels [physout, right, left]);  P_i = C_i Y_kept.
3. Site 0 = A_0 P_1 (labels [right, physout, left]).

Precision: "fp64" runs every contraction in FP64 (the algorithmic change
alone); "fp32" / "bf16x3" run the contractions on FP32 / split-bf16 MFMA with
FP32 accumulation, while K, the Jacobi rotations and the output isometries
stay FP64 (the eigensolver then stops at orthogonality 1e-10 instead of
n * eps: K formed in FP32 carries ~1e-7 relative rounding anyway).  Flop
count per bulk site at chi=512, D=5, d=2: ~52 GFLOP versus ~250 GFLOP for
the reference's QR + SVD on the 2560-bond chain.
"""
from __future__ import annotations

import numpy as np
import torch

from tncontract_amd import _lib as L
from tncontract_amd import tensor as tsr
from tncontract_amd.onedim import onedim_core as core

__all__ = ["apply_mpo_compress", "SweepStats"]


class SweepStats:
    """Per-bond record of one structured sweep (for reports and the rank model)."""

    def __init__(self):
        self.spectra = {}   # bond i -> sigma / sigma_0 (host, descending)
        self.kept = {}      # bond i -> kept rank
        self.avail = {}     # bond i -> number of admissible singular values
        self.sweeps = {}    # bond i -> Jacobi sweeps of the eigensolver


def _canon(t, labels):
    """Strided view of t.data with axes in the order of `labels`."""
    return t.data.permute([t.labels.index(x) for x in labels])


def _ones(shape, dtype):
    host = np.ones(shape, dtype=np.float64 if dtype == torch.float64 else np.float32)
    return torch.from_numpy(host).to(L.device())


def _reorthonormalise(Y):
    """Y (n x k, FP64, orthonormal columns up to a small error e) ->
    Y (3 I - Y^T Y) / 2, orthonormal up to O(e^2) (Newton-Schulz step of the
    polar factor)."""
    S = tsr._contract_data(Y, [0], Y, [0])                # Y^T Y
    eye = L.diag_embed(_ones((Y.shape[1],), torch.float64))
    X = L.axpby(S, eye, -0.5, 1.5)
    return tsr._contract_data(Y, [1], X, [0])


def _cdt(x, dtype):
    return L.convert(x, dtype) if x.dtype != dtype else x


# Jacobi orthogonality tolerance of the eigensolver when K is formed from
# FP32 / bf16x3 contractions: K itself then carries ~1e-7 relative rounding,
# so converging the rotations to n * eps buys nothing; 1e-10 keeps the
# eigensolver's share of the state error far below the contractions' share.
MIXED_EIG_TOL = 1e-10


def apply_mpo_compress(mps, mpo, chi=None, threshold=1e-15, precision="fp64", stats=None,
                       eig_tol=None):
    if len(mps) != len(mpo):
        raise ValueError("MPS and MPO have different lengths")
    N = len(mps)
    work = torch.float64 if precision in ("fp64", "native") else torch.float32
    prec = "native" if work == torch.float64 else precision
    lab_B = [mps.phys_label, mps.left_label, mps.right_label]
    lab_W = [mpo.left_label, mpo.right_label, mpo.physout_label, mpo.physin_label]
    B = [_cdt(_canon(t, lab_B), work) for t in mps]      # (s, a, b)
    W = [_cdt(_canon(t, lab_W), work) for t in mpo]      # (alpha, beta, t, s)
    d = [w.shape[2] for w in W]
    for i in range(N):
        if B[i].shape[0] != W[i].shape[3]:
            raise ValueError("physical dimension mismatch at site %d" % i)
    con = lambda x, ax1, y, ax2: tsr._contract_data(x, ax1, y, ax2, prec)  # noqa: E731
    if eig_tol is None:
        eig_tol = 0.0 if work == torch.float64 else MIXED_EIG_TOL   # 0: n * eps

    # ---- 1. left environments E_i (a, alpha, a', alpha'), normalised
    E = [None] * N
    E[0] = _ones((1, 1, 1, 1), work)
    for i in range(N - 1):
        Wt = con(W[i], [2], W[i], [2])                    # (al, be, s, al', be', s')
        T1 = con(E[i], [0], B[i], [1])                    # (al, a', al', s, b)
        T2 = con(T1, [0, 2, 3], Wt, [0, 3, 2])            # (a', b, be, be', s')
        En = con(T2, [0, 4], B[i], [1, 0])                # (b, be, be', b')
        L.normalize_(En)
        E[i + 1] = En.permute(0, 1, 3, 2)                 # (b, be, b', be') view

    # bond dimensions the reference's QR left sweep produces (min(rows, cols))
    qb = [1]
    for i in range(N - 1):
        qb.append(min(d[i] * qb[-1], B[i].shape[2] * W[i].shape[1]))

    # ---- 2. right-to-left truncation
    sites = [None] * N
    P = _ones((1, 1, 1), work)                            # (b, beta, k)
    for i in range(N - 1, 0, -1):
        U1 = con(B[i], [2], P, [0])                       # (s, a, be, k)
        C = con(U1, [2, 0], W[i], [1, 3])                 # (a, k, al, t)
        F = con(E[i], [0, 1], C, [0, 2])                  # (a', al', k, t)
        Kt = con(C, [0, 2], F, [0, 1])                    # (k', t', k, t)
        r = C.shape[1]
        n = d[i] * r
        # rows (t', k'), cols (t, k); an FP32 K goes in as FP32 so that the
        # eigensolver's shift covers its rounding
        Km = Kt.permute(1, 0, 3, 2)
        info = {}
        u, lam = tsr._eigh_psd_device(Km, 2, eig_tol, info)
        lam_h = lam.cpu().numpy()
        avail = min(n, qb[i])
        sig = np.sqrt(np.maximum(lam_h[:avail], 0.0))
        if sig.size == 0 or sig[0] == 0.0:
            raise NotImplementedError("zero state in the structured sweep")
        rel = sig / sig[0]
        keep = int(np.count_nonzero(rel > threshold))
        if chi:
            keep = min(keep, chi)
        if stats is not None:
            stats.spectra[i] = rel
            stats.kept[i] = keep
            stats.avail[i] = avail
            stats.sweeps[i] = info["sweeps"]
        if eig_tol > 0.0:
            # the rotations stopped at orthogonality eig_tol: one Newton-Schulz
            # step Y (3 I - Y^T Y) / 2 restores an isometry to rounding
            Y2 = _reorthonormalise(u[:, :keep])
            Y = Y2.view(d[i], r, keep)                    # (t, k, kept), FP64 isometry
        else:
            Y = u.view(d[i], r, n)[:, :, :keep]           # (t, k, kept), FP64 isometry
        sites[i] = tsr.Tensor._wrap(L.contiguous(Y), [mpo.physout_label, mps.right_label,
                                                       mps.left_label])
        P = con(C, [3, 1], _cdt(Y, work), [0, 1])         # (a, al, kept)
    U1 = con(B[0], [2], P, [0])                           # (s, a=1, be=1, k)
    C = con(U1, [2, 0], W[0], [1, 3])                     # (a=1, k, al=1, t)
    s0 = L.materialize(_cdt(C, torch.float64).permute(1, 3, 0, 2)).view(C.shape[1], d[0], 1)
    sites[0] = tsr.Tensor._wrap(s0,
                                [mps.right_label, mpo.physout_label, mps.left_label])
    return core.MatrixProductState._from_owned(sites, mps.left_label, mps.right_label,
                                               mpo.physout_label)

