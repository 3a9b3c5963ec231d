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

Resolution: K holds sigma^2, so a bond's eigenvalues are resolved only down
to K's rounding level: ASYM_C 2 ||K - K^T||_F / sqrt(n) (the asymmetry of
the formed K samples the rounding of its two triangles, and rounding is
close to random noise, whose spectral norm is ~2 / sqrt(n) of its Frobenius
norm), or the eigenvalues of K's structural null space where the bond's rank
bound leaves one, and at least eps sqrt(n) lambda_0.  Below sigma / sigma_0 = f = sqrt(FLOOR_C level /
lambda_0) a value cannot be told from rounding (f ~ 1e-7 for an FP64 K,
~1e-3 .. 1e-2 for an FP32 one).  A bond's rank is therefore taken here only
when it is decided above f: by a threshold >= f, or by the chi / QR-bond cap
binding above f.  A bond whose rank would be decided below f (a threshold-
bound truncation of a decaying or rank-deficient spectrum, a chi cut deep in
the tail, a zero state) sends the whole call to the exact path,
contract_mps_mpo + svd_compress, so the result always has the reference's
bond dimensions; ``SweepStats.fallback`` records the bond that did.

Precision: "fp64" runs every contraction in FP64 (the algorithmic change
alone); "fp32" / "bf16x3" run the contractions on FP32 / split-bf16 MFMA with
FP32 accumulation (bf16x3: each FP32 operand split into two bf16 halves,
three products, ~2^-16 relative rounding per product), while K, the Jacobi
rotations and the output isometries stay FP64 (the eigensolver then stops at orthogonality 1e-8 instead of
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
        self.floor = {}     # bond i -> resolution floor f of sigma / sigma_0
        self.fallback = None  # bond whose rank sent the call to the exact path


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
# so converging the rotations to n * eps buys nothing.  Measured on the XXZ
# chain (N=64, chi=128 and N=22, chi=512, FP32 contractions): state errors
# against the oracle 9.7e-12 / 7.0e-12 at 1e-10, 6.9e-12 / 7.7e-12 at 1e-8,
# 1.0e-10 / 1.2e-11 at 1e-6.
MIXED_EIG_TOL = 1e-8


# lower bound of K's rounding level, per unit eps_k sqrt(n) lambda_0, by the
# precision the contractions accumulate in.  bf16x3 accumulates in FP32; its
# products carry more (each FP32 operand split into two bf16 halves, 16
# significant bits, the lo * lo product dropped: ~2^-16 of |x y|), but that
# rounding differs between K's two triangles and is measured by the
# asymmetry term (XXZ chain N=64, chi=128: ||K - K^T|| level 2.6e-5 lambda_0
# at a 256-wide bond, K's null-space eigenvalues 5e-7 lambda_0, where
# 2^-16 sqrt(n) would claim 2.4e-4)
_EPS_K = {"fp64": 2.0 ** -52, "native": 2.0 ** -52, "fp32": 2.0 ** -23, "bf16x3": 2.0 ** -23}
# the rounding of K, estimated from its asymmetry ||K - K^T||_F (the two
# triangles are separate sums; the environment's own rounding, which is
# symmetric, is covered by the factor ASYM_C), or observed directly in the
# eigenvalues of K's structural null space; eigenvalues below FLOOR_C times
# that level are taken as unresolved
ASYM_C = 8.0
FLOOR_C = 4.0


class _Unresolved(Exception):
    """A bond's rank would be decided below the resolution floor."""

    def __init__(self, bond):
        super().__init__(bond)
        self.bond = bond


def bond_rank(lam, n, avail, chi, threshold, eps_k, asym=0.0):
    """Kept rank at a bond from the eigenvalues lam (descending, host) of its
    Gram matrix K, by the reference's rule (sigma / sigma_0 > threshold, then
    at most chi; at most `avail`, the bond dimension the reference's QR sweep
    gives, which bounds the rank of K), or None when that rank is decided
    below the resolution floor.  asym: ||K - K^T||_F of the K the eigensolver
    symmetrised.  The rounding level of K (a spectral norm) is the largest
    of ASYM_C * 2 asym / sqrt(n) (rounding errors are close to random, and a
    random n x n matrix has spectral norm ~2 / sqrt(n) of its Frobenius
    norm), the magnitude of the eigenvalues past `avail` (zero in exact
    arithmetic) and eps_k sqrt(n) lambda_0; the floor is f = sqrt(FLOOR_C
    level / lambda_0) in sigma / sigma_0.  Returns (keep or None, rel sigma
    / sigma_0 over avail, f)."""
    lam = np.asarray(lam, dtype=np.float64)
    if lam.size == 0 or not np.isfinite(lam[0]) or lam[0] <= 0.0 or not np.all(np.isfinite(lam)):
        return None, np.zeros(0), float("inf")
    lam0 = float(lam[0])
    level = max(ASYM_C * 2.0 * float(asym) / np.sqrt(n), eps_k * np.sqrt(n) * lam0)
    if avail < lam.size:
        level = max(level, float(np.max(np.abs(lam[avail:]))))
    f = float(np.sqrt(FLOOR_C * level / lam0))
    rel = np.sqrt(np.maximum(lam[:avail], 0.0) / lam0)
    cap = min(chi, avail) if chi else avail
    if threshold >= f:
        return min(int(np.count_nonzero(rel > threshold)), cap), rel, f
    if int(np.count_nonzero(rel > f)) >= cap:
        return cap, rel, f
    return None, rel, f


def apply_mpo_compress(mps, mpo, chi=None, threshold=1e-15, precision="fp64", stats=None,
                       eig_tol=None):
    """contract_mps_mpo(mps, mpo) followed by svd_compress(chi, threshold)
    (onedim_core.py:1691, :463) by the structured sweep; falls back to exactly
    those two calls when a bond's rank is not resolvable by the sweep."""
    try:
        return _apply_mpo_compress(mps, mpo, chi, threshold, precision, stats, eig_tol)
    except _Unresolved as e:
        if stats is not None:
            stats.fallback = e.bond
        out = core.contract_mps_mpo(mps, mpo)
        out.svd_compress(chi=chi, threshold=threshold)
        return out


def _apply_mpo_compress(mps, mpo, chi, threshold, precision, stats, eig_tol):
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
        try:
            u, lam = tsr._eigh_psd_device(Km, 2, eig_tol, info)
        except np.linalg.LinAlgError:
            # no convergence (a zero or non-finite K among them): the exact
            # path decides, with its own retry and diagnostics
            raise _Unresolved(i)
        lam_h = lam.cpu().numpy()
        avail = min(n, qb[i])
        # ||K - K^T||_F: K is the (t', k') x (t, k) view, K^T swaps the groups
        asym = float(np.sqrt(L.sumsq(L.axpby(Km, Km.permute(2, 3, 0, 1), 1.0, -1.0))))
        keep, rel, f = bond_rank(lam_h, n, avail, chi, threshold, _EPS_K[precision], asym)
        if stats is not None:
            stats.spectra[i] = rel
            stats.avail[i] = avail
            stats.sweeps[i] = info["sweeps"]
            stats.floor[i] = f
        if keep is None:
            raise _Unresolved(i)
        if stats is not None:
            stats.kept[i] = keep
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

