"""MPS/MPO constructors and local expectation values (tncontract/onedim/onedim_utils.py),
plus the seeded generators and the XXZ MPO used by the benchmark workloads.
"""
from __future__ import annotations

import numpy as np

from tncontract_amd import tensor as tnc
from tncontract_amd.onedim import onedim_core as onedim

__all__ = ["init_mps_random", "init_mps_allzero", "init_mps_logical", "onebody_sum_mpo",
           "expvals_mps", "xxz_mpo", "xxz_bulk_w", "random_mps_arrays", "random_mps",
           "decaying_mps_arrays"]


def init_mps_random(nsites, physdim, bonddim=1, left_label="left", right_label="right",
                    phys_label="phys"):
    """onedim_utils.py:23-60: np.random.rand entries, each site rescaled so its
    largest singular value (over [phys,left] | right) is 1."""
    if not np.iterable(physdim):
        physdim = [physdim] * nsites
    if not np.iterable(bonddim):
        bonddim = [bonddim] * (nsites - 1)
    bonddim = [1] + list(bonddim) + [1]
    tensors = []
    for i in range(nsites):
        rt = tnc.Tensor(np.random.rand(physdim[i], bonddim[i], bonddim[i + 1]),
                        [phys_label, left_label, right_label])
        U, S, V, s = tnc._tensor_svd(rt, [phys_label, left_label])
        s0 = float(s[0].item())
        S.scale_(1.0 / s0)
        rt = U["svd_in", ] * S["svd_out", ]
        rt = rt["svd_in", ] * V["svd_out", ]
        tensors.append(rt)
    return onedim.MatrixProductState(tensors, left_label=left_label, right_label=right_label,
                                     phys_label=phys_label)


def init_mps_allzero(nsites, physdim, left_label="left", right_label="right", phys_label="phys"):
    """onedim_utils.py:63-87"""
    if not np.iterable(physdim):
        physdim = [physdim] * nsites
    tensors = []
    for j in range(nsites):
        t = np.zeros(physdim[j])
        t[0] = 1.0
        tensors.append(tnc.Tensor(t.reshape(physdim[j], 1, 1), [phys_label, left_label, right_label]))
    return onedim.MatrixProductState(tensors, left_label=left_label, right_label=right_label,
                                     phys_label=phys_label)


def init_mps_logical(nsites, basis_state, physdim, left_label="left", right_label="right",
                     phys_label="phys"):
    """onedim_utils.py:90-119"""
    if not np.iterable(physdim):
        physdim = [physdim] * nsites
    if not np.iterable(basis_state):
        basis_state = [basis_state] * nsites
    tensors = []
    for j in range(nsites):
        t = np.zeros(physdim[j])
        t[basis_state[j]] = 1.0
        tensors.append(tnc.Tensor(t.reshape(physdim[j], 1, 1), [phys_label, left_label, right_label]))
    return onedim.MatrixProductState(tensors, left_label=left_label, right_label=right_label,
                                     phys_label=phys_label)


def onebody_sum_mpo(terms, output_label=None):
    """onedim_utils.py:122-172 (real terms only: the GPU path is real)."""
    tensors = []
    mats = []
    for term1 in terms:
        if isinstance(term1, tnc.Tensor):
            term = term1.copy()
            if output_label is not None:
                term.move_index(output_label, 0)
            mats.append(term.numpy())
        else:
            mats.append(np.asarray(term1))
    for i, term in enumerate(mats):
        if np.iscomplexobj(term):
            raise NotImplementedError("complex MPO terms are not supported on the GPU path")
        d0, d1 = term.shape
        eye = np.eye(d0, d1)
        if i == 0:
            B = np.stack([term, eye], axis=-1)
            tensors.append(tnc.Tensor(B, ["physout", "physin", "right"]))
        elif i == len(mats) - 1:
            B = np.stack([eye, term], axis=-1)
            tensors.append(tnc.Tensor(B, ["physout", "physin", "left"]))
        else:
            B = np.zeros((d0, d1, 2, 2))
            B[:, :, 0, 0] = eye
            B[:, :, 1, 0] = term
            B[:, :, 1, 1] = eye
            tensors.append(tnc.Tensor(B, ["physout", "physin", "left", "right"]))
    return onedim.MatrixProductOperator(tensors, left_label="left", right_label="right",
                                        physin_label="physin", physout_label="physout")


def expvals_mps(mps, oplist=[], sites=None, output_label=None, canonised=None):
    """onedim_utils.py:175-255.  Returns a complex numpy array like the reference."""
    if sites is None:
        sites = range(len(mps))
    if not np.iterable(sites):
        sites = [sites]
    N = len(sites)
    expvals = np.zeros(N, dtype=complex)
    oplist_new = oplist if isinstance(oplist, list) else [oplist] * N
    if canonised == "left":
        mps.reverse()
        oplist_new = oplist_new[::-1]
    elif canonised != "right":
        mps.right_canonise()
    center = 0
    for i, site in enumerate(sites):
        mps.left_canonise(center, site)
        center = site
        op = oplist_new[i]
        if not isinstance(op, tnc.Tensor):
            op = tnc.Tensor(op)
        A = mps[site]
        if output_label is None:
            out_label, in_label = op.labels[0], op.labels[1]
        else:
            out_label = output_label
            in_label = [x for x in op.labels if x is not out_label][0]
        exp = tnc.contract(A, op, mps.phys_label, in_label)
        exp = tnc.contract(A, exp, mps.phys_label, out_label)
        exp.contract_internal(mps.left_label, mps.left_label, index1=0, index2=1)
        exp.contract_internal(mps.right_label, mps.right_label, index1=0, index2=1)
        expvals[i] = exp.numpy()
    if canonised == "left":
        mps.reverse()
        expvals = expvals[::-1]
    return expvals


# ---------------------------------------------------------------- workload builders
def xxz_bulk_w(Jz=1.0, Jxy=1.0, hz=0.5):
    """W[a, b, out, in] of H = sum_i Jz Sz_i Sz_{i+1} + Jxy/2 (S+_i S-_{i+1} +
    S-_i S+_{i+1}) + hz sum_i Sz_i, spin-1/2 (S = sigma/2), bond D = 5.
    Row a = 4 is the left boundary, column b = 0 the right boundary."""
    sz = np.array([[0.5, 0.0], [0.0, -0.5]])
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    sm = sp.T.copy()
    eye = np.eye(2)
    W = np.zeros((5, 5, 2, 2))
    W[0, 0], W[1, 0], W[2, 0], W[3, 0] = eye, sp, sm, sz
    W[4, 0], W[4, 1], W[4, 2], W[4, 3], W[4, 4] = hz * sz, 0.5 * Jxy * sm, 0.5 * Jxy * sp, Jz * sz, eye
    return W


def xxz_mpo(N, Jz=1.0, Jxy=1.0, hz=0.5):
    """XXZ chain MPO (bond D=5) in the reference MPO format (labels
    left/right/physout/physin).  The reference has no XXZ builder."""
    W = xxz_bulk_w(Jz, Jxy, hz)
    tensors = []
    for i in range(N):
        if N == 1:
            tensors.append(tnc.Tensor(W[4, 0], ["physout", "physin"]))
        elif i == 0:
            tensors.append(tnc.Tensor(W[4, :], ["right", "physout", "physin"]))
        elif i == N - 1:
            tensors.append(tnc.Tensor(W[:, 0], ["left", "physout", "physin"]))
        else:
            tensors.append(tnc.Tensor(W, ["left", "right", "physout", "physin"]))
    return onedim.MatrixProductOperator(tensors, "left", "right", "physout", "physin")


def random_mps_arrays(N, d, chi, seed, dtype=np.float64):
    """Seeded random MPS site arrays (phys, left, right): bonds [1, chi, ...,
    chi, 1], entries i.i.d. N(0,1) from numpy.random.default_rng(seed), then
    every site divided by ||psi||^(1/N) so the state is normalised.  The norm
    is computed on the host with per-site rescaled transfer matrices (exact
    in FP64, no overflow)."""
    rng = np.random.default_rng(seed)
    arrs = []
    for i in range(N):
        cl = 1 if i == 0 else chi
        cr = 1 if i == N - 1 else chi
        arrs.append(rng.standard_normal((d, cl, cr)))
    # log ||psi||^2 by the transfer-matrix recursion E <- sum_s A_s^T E A_s
    E = np.ones((1, 1))
    log2 = 0.0
    for a in arrs:
        E = np.einsum("ab,sac,sbd->cd", E, a, a, optimize=True)
        s = np.abs(E).max()
        E /= s
        log2 += np.log(s)
    log2 += np.log(E[0, 0])
    fac = np.exp(-0.5 * log2 / N)
    return [(a * fac).astype(dtype) for a in arrs]


def random_mps(N, d, chi, seed, phys_label="phys", left_label="left", right_label="right"):
    arrs = random_mps_arrays(N, d, chi, seed)
    return onedim.MatrixProductState([tnc.Tensor(a, [phys_label, left_label, right_label])
                                      for a in arrs], left_label, right_label, phys_label)


def decaying_mps_arrays(N, d, chi, taus, seed, dtype=np.float32):
    """Seeded MPS site arrays (phys, left, right) with exponentially decaying
    bond spectra: bonds b_i = min(d^i, d^(N-i), chi); site i is a random right
    isometry (rows of a QR of a Gaussian, b_i x d b_{i+1}) whose left index k
    is scaled by exp(-k / taus[i - 1]) (site 0 unscaled).  The Schmidt values
    at bond i then fall off roughly like exp(-k / taus[i - 1]).  Returned in
    `dtype` (the learned-rank workload's data are FP32)."""
    rng = np.random.default_rng(seed)
    b = [min(d ** i, d ** (N - i), chi) for i in range(N + 1)]
    arrs = []
    for i in range(N):
        l, r = b[i], b[i + 1]
        q, _ = np.linalg.qr(rng.standard_normal((d * r, l)))     # (d r) x l, orthonormal columns
        site = q.T.reshape(l, d, r).transpose(1, 0, 2)           # (phys, left, right)
        if i > 0:
            site = site * np.exp(-np.arange(l) / float(taus[i - 1]))[None, :, None]
        arrs.append(np.ascontiguousarray(site).astype(dtype))
    return arrs
