"""Batched uneven contraction + truncated SVD (many independent jobs).

``batched_truncate(lefts, rights, rel_threshold, chi)`` computes, for every
job j, C_j = L_j @ R_j and its truncated SVD (keep sigma/sigma_0 >
rel_threshold, at most chi), all jobs in shared launches of
libtncontract_amd (grouped GEMM of the FP32 factors with FP64 accumulation
and autotuned tile shapes, batched one-sided Jacobi with FP64 rotations,
replayed as a captured HIP graph).  Per job this is what the reference
computes with ``contract`` (tensor.py:635) followed by
``truncated_svd(..., absolute=False)`` (tensor.py:1112); the product is
FP64-accurate, so the kept rank is the one the reference's FP64 tensordot of
the same factors gives.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from tncontract_amd import _lib as L


class tn_trunc_job(ctypes.Structure):
    _fields_ = [("L", ctypes.c_void_p), ("R", ctypes.c_void_p), ("m", ctypes.c_int64),
                ("k", ctypes.c_int64), ("n", ctypes.c_int64), ("U", ctypes.c_void_p),
                ("S", ctypes.c_void_p), ("Vt", ctypes.c_void_p)]


class tn_svd_job(ctypes.Structure):
    _fields_ = [("A", ctypes.c_void_p), ("m", ctypes.c_int64), ("n", ctypes.c_int64),
                ("U", ctypes.c_void_p), ("S", ctypes.c_void_p), ("Vt", ctypes.c_void_p)]


def _bind():
    lib = L.lib()
    if not hasattr(lib, "_tn_batched_bound"):
        sp = ctypes.POINTER(tn_svd_job)
        lib.tn_batched_svd_workspace.restype = ctypes.c_size_t
        lib.tn_batched_svd_workspace.argtypes = [sp, ctypes.c_int]
        lib.tn_batched_svd.restype = ctypes.c_int
        lib.tn_batched_svd.argtypes = [sp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                       ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        jp = ctypes.POINTER(tn_trunc_job)
        lib.tn_batched_truncate_workspace.restype = ctypes.c_size_t
        lib.tn_batched_truncate_workspace.argtypes = [jp, ctypes.c_int]
        lib.tn_batched_truncate.restype = ctypes.c_int
        lib.tn_batched_truncate.argtypes = [jp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                            ctypes.c_size_t, ctypes.c_void_p]
        lib._tn_batched_bound = True
    return lib


def batched_truncate(lefts, rights, rel_threshold=1e-8, chi=0, max_sweeps=60):
    """lefts[j]: (m_j, k_j) FP32 device tensor, rights[j]: (k_j, n_j).
    Returns (U, S, Vt, kept, sweeps): per-job full thin factors (device) and
    host arrays of kept ranks and Jacobi sweeps."""
    lib = _bind()
    nj = len(lefts)
    jobs = (tn_trunc_job * nj)()
    outs = []
    keep_alive = []
    for j, (a, b) in enumerate(zip(lefts, rights)):
        a = L.contiguous(L.convert(a, torch.float32))
        b = L.contiguous(L.convert(b, torch.float32))
        m, k = a.shape
        k2, n = b.shape
        if k != k2:
            raise ValueError("job %d: inner dimensions differ (%d vs %d)" % (j, k, k2))
        r = min(m, n)
        U = torch.empty((m, r), dtype=torch.float32, device=a.device)
        S = torch.empty((r,), dtype=torch.float64, device=a.device)
        Vt = torch.empty((r, n), dtype=torch.float32, device=a.device)
        jobs[j] = tn_trunc_job(a.data_ptr(), b.data_ptr(), m, k, n, U.data_ptr(), S.data_ptr(),
                               Vt.data_ptr())
        outs.append((U, S, Vt))
        keep_alive += [a, b]
    ws = L.workspace(lib.tn_batched_truncate_workspace(jobs, nj))
    kept = (ctypes.c_int * nj)()
    sweeps = (ctypes.c_int * nj)()
    L.check(lib.tn_batched_truncate(jobs, nj, float(rel_threshold), int(chi or 0), int(max_sweeps),
                                    kept, sweeps, L.ptr(ws), ws.numel(), L.stream_ptr()))
    return ([o[0] for o in outs], [o[1] for o in outs], [o[2] for o in outs],
            np.array(kept[:]), np.array(sweeps[:]))


def batched_svd(mats, max_sweeps=40):
    """Thin FP64 SVDs of many device matrices (np.linalg.svd per matrix,
    tensor.py:915) in shared launches.  mats[j]: (m_j, n_j) device tensor.
    Returns (U, S, Vt, sweeps): per-job device factors (S descending) and the
    host array of Jacobi sweeps."""
    lib = _bind()
    nj = len(mats)
    jobs = (tn_svd_job * nj)()
    outs, keep_alive = [], []
    for j, a in enumerate(mats):
        a = L.contiguous(L.convert(a, torch.float64))
        m, n = a.shape
        r = min(m, n)
        U = torch.empty((m, r), dtype=torch.float64, device=a.device)
        S = torch.empty((r,), dtype=torch.float64, device=a.device)
        Vt = torch.empty((r, n), dtype=torch.float64, device=a.device)
        jobs[j] = tn_svd_job(a.data_ptr(), m, n, U.data_ptr(), S.data_ptr(), Vt.data_ptr())
        outs.append((U, S, Vt))
        keep_alive.append(a)
    if nj == 0:
        return [], [], [], np.zeros(0, dtype=int)
    ws = L.workspace(lib.tn_batched_svd_workspace(jobs, nj))
    sweeps = (ctypes.c_int * nj)()
    L.check(lib.tn_batched_svd(jobs, nj, int(max_sweeps), sweeps, L.ptr(ws), ws.numel(),
                               L.stream_ptr()))
    return ([o[0] for o in outs], [o[1] for o in outs], [o[2] for o in outs],
            np.array(sweeps[:]))


def config3_jobs(njobs, seed, chi_lo=16, chi_hi=512, d=2):
    """Seeded jobs of the batched-truncation workload: L_j (chi_l d x chi_m)
    and R_j (chi_m x d chi_r) with L_j R_j = U diag(sigma) V^T, U, V random
    orthogonal (QR of Gaussians), sigma_k = exp(-k / tau), tau ~ U[4, 64].
    Returns host float32 arrays and the exact spectra."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(njobs):
        cl, cm, cr = (int(x) for x in rng.integers(chi_lo, chi_hi + 1, size=3))
        tau = float(rng.uniform(4.0, 64.0))
        m, n = cl * d, d * cr
        r = min(cm, m, n)
        u, _ = np.linalg.qr(rng.standard_normal((m, r)))
        v, _ = np.linalg.qr(rng.standard_normal((n, r)))
        x, _ = np.linalg.qr(rng.standard_normal((cm, r)))  # cm x r, orthonormal columns
        sig = np.exp(-np.arange(r) / tau)
        Lm = ((u * sig) @ x.T).astype(np.float32)          # m x cm
        Rm = (x @ v.T).astype(np.float32)                   # cm x n
        out.append((Lm, Rm, sig))
    return out
