"""Oracle-side builders and dense brute-force helpers.

TEST INFRASTRUCTURE ONLY (same import rule as ``oracle/tn_oracle.py``).

* ``xxz_mpo``: the XXZ chain MPO (bond D=5) in the reference's MPO format
  (labels left/right/physout/physin, onedim_core.py:1329-1367).  The reference
  ships no XXZ builder (its only MPO builder is onedim_utils.py:122-172); this
  one is pinned against ``xxz_dense_hamiltonian`` (an independent Kronecker-sum
  construction) in tests/test_oracle.py.
* ``mps_to_dense`` / ``mpo_to_dense``: contract a chain into a dense vector /
  matrix (used for brute-force checks at N <= 12).
* ``tt_truncate_dense``: sequential right-to-left truncated SVD of a dense
  state vector -- the closed-form answer ``svd_compress`` must reproduce.
"""
from __future__ import annotations

import numpy as np

from oracle.tn_oracle import OMPO, OMPS, OTensor

SZ = np.array([[0.5, 0.0], [0.0, -0.5]])
SP = np.array([[0.0, 1.0], [0.0, 0.0]])
SM = np.array([[0.0, 0.0], [1.0, 0.0]])
I2 = np.eye(2)


def xxz_bulk_w(Jz=1.0, Jxy=1.0, hz=0.5):
    """W[a, b, out, in] of H = sum Jz Sz Sz + Jxy/2 (S+S- + S-S+) + hz sum Sz.
    Row index a=4 is the left boundary, column b=0 the right boundary."""
    W = np.zeros((5, 5, 2, 2))
    W[0, 0] = I2
    W[1, 0] = SP
    W[2, 0] = SM
    W[3, 0] = SZ
    W[4, 0] = hz * SZ
    W[4, 1] = 0.5 * Jxy * SM
    W[4, 2] = 0.5 * Jxy * SP
    W[4, 3] = Jz * SZ
    W[4, 4] = I2
    return W


def xxz_mpo(N, Jz=1.0, Jxy=1.0, hz=0.5):
    """Oracle XXZ MPO; edge tensors lack the outer bond (a dummy index is
    added by the container, as onedim_core.py:59-61 does)."""
    W = xxz_bulk_w(Jz, Jxy, hz)
    tensors = []
    for i in range(N):
        if N == 1:
            t = OTensor(W[4, 0], ["physout", "physin"])
        elif i == 0:
            t = OTensor(W[4, :], ["right", "physout", "physin"])
        elif i == N - 1:
            t = OTensor(W[:, 0], ["left", "physout", "physin"])
        else:
            t = OTensor(W, ["left", "right", "physout", "physin"])
        tensors.append(t)
    return OMPO(tensors, "left", "right", "physout", "physin")


def _site_op(op, i, N):
    out = np.array([[1.0]])
    for j in range(N):
        out = np.kron(out, op if j == i else I2)
    return out


def xxz_dense_hamiltonian(N, Jz=1.0, Jxy=1.0, hz=0.5):
    """Independent Kronecker-sum construction (site 0 = most significant)."""
    H = np.zeros((2 ** N, 2 ** N))
    for i in range(N - 1):
        H += Jz * _site_op(SZ, i, N) @ _site_op(SZ, i + 1, N)
        H += 0.5 * Jxy * (_site_op(SP, i, N) @ _site_op(SM, i + 1, N)
                          + _site_op(SM, i, N) @ _site_op(SP, i + 1, N))
    for i in range(N):
        H += hz * _site_op(SZ, i, N)
    return H


def _site_array(t, phys, left, right):
    """Return site data ordered (left, phys, right)."""
    return np.transpose(t.data, [t.labels.index(left), t.labels.index(phys),
                                 t.labels.index(right)])


def mps_to_dense(mps):
    """Dense state vector, site 0 most significant."""
    v = np.ones((1, 1))
    for t in mps:
        a = _site_array(t, mps.phys_label, mps.left_label, mps.right_label)
        v = np.tensordot(v, a, ([1], [0]))  # (rows, phys, right)
        v = v.reshape(-1, a.shape[2])
    assert v.shape[1] == 1
    return v[:, 0]


def mpo_to_dense(mpo):
    M = np.ones((1, 1, 1))  # (out, in, right)
    for t in mpo:
        a = np.transpose(t.data, [t.labels.index(mpo.left_label), t.labels.index(mpo.physout_label),
                                  t.labels.index(mpo.physin_label), t.labels.index(mpo.right_label)])
        M = np.einsum("oir,rabs->oaibs", M, a)
        s = M.shape
        M = M.reshape(s[0] * s[1], s[2] * s[3], s[4])
    return M[:, :, 0]


def dense_to_mps_arrays(v, N, d):
    """Exact (untruncated) left-to-right SVD split of a dense vector into
    (phys, left, right) site arrays."""
    arrs = []
    rest = v.reshape(1, -1)
    for i in range(N - 1):
        chi_l = rest.shape[0]
        m = rest.reshape(chi_l * d, -1)
        u, s, vt = np.linalg.svd(m, full_matrices=False)
        arrs.append(np.transpose(u.reshape(chi_l, d, -1), (1, 0, 2)))
        rest = s[:, None] * vt
    arrs.append(np.transpose(rest.reshape(rest.shape[0], d, 1), (1, 0, 2)))
    return arrs


def tt_truncate_dense(v, N, d, chi, rel_threshold=0.0):
    """Sequential right-to-left truncated SVD of a dense vector (the TT
    rounding that svd_compress performs once the left part is an isometry).
    Returns the truncated dense vector."""
    # right-to-left: at each bond keep the leading chi singular triplets
    rest = v.reshape(-1, 1)  # (left_dims, kept_right)
    right_factor = np.ones((1, 1))  # (kept, right-dense)
    right_dense = np.ones((1, 1))
    for i in range(N - 1, 0, -1):
        m = rest.reshape(-1, d * rest.shape[1])  # (d^i, d*r)
        u, s, vt = np.linalg.svd(m, full_matrices=False)
        keep = s / s[0] > rel_threshold
        k = int(np.count_nonzero(keep))
        if chi:
            k = min(k, chi)
        vt = vt[:k]
        # new right block: (k, d*r) x (r, dense_right) -> (k, d * dense_right)
        vr = vt.reshape(k, d, rest.shape[1])
        right_dense = np.tensordot(vr, right_dense, ([2], [0])).reshape(k, -1)
        rest = m @ vt.T  # (d^i, k)
    return (rest @ right_dense).reshape(-1)
