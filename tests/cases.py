"""Seeded inputs shared by the oracle and GPU tests: states with known
Schmidt spectra and simple MPOs, as plain NumPy arrays."""
import numpy as np


def schmidt_pair(sigma, d, seed):
    """Two-site MPS psi[s, t] = sum_b U[s, b] sigma_b V[t, b] with orthogonal
    U, V (d x d): its only bond has Schmidt values sigma (descending).
    Returns (site arrays (phys, left, right), U[:, :k], V[:, :k], dense)."""
    r = np.random.default_rng(seed)
    U = np.linalg.qr(r.standard_normal((d, d)))[0]
    V = np.linalg.qr(r.standard_normal((d, d)))[0]
    k = len(sigma)
    s0 = (U[:, :k] * sigma)[:, None, :]          # (phys, left=1, right=k)
    s1 = V[:, :k][:, :, None]                     # (phys, left=k, right=1)
    dense = (U[:, :k] * sigma) @ V[:, :k].T
    return [s0, s1], U[:, :k], V[:, :k], dense


def _isometry(rows, cols, rng):
    q, _ = np.linalg.qr(rng.standard_normal((rows, cols)))
    return q


def schmidt_chain(N, d, sigma, seed):
    """Site arrays (phys, left, right) of an N-site MPS whose bond between
    sites N//2 - 1 and N//2 has exactly the Schmidt values `sigma`: the left
    half is a random left isometry (bonds min(d^i, ...) growing from 1), the
    right half a random right isometry, and diag(sigma) sits on the middle
    site's right bond.  len(sigma) must not exceed d^(N//2) nor d^(N - N//2)."""
    rng = np.random.default_rng(seed)
    k = len(sigma)
    h = N // 2
    assert k <= d ** h and k <= d ** (N - h)
    lb = [1]
    for i in range(h):
        lb.append(min(d * lb[-1], k if i == h - 1 else d ** (i + 1)))
    rb = [1]
    for i in range(N - h):
        rb.append(min(d * rb[-1], k if i == N - h - 1 else d ** (i + 1)))
    rb = rb[::-1]                      # bonds of the right half, left to right
    arrs = []
    for i in range(h):                 # left isometries: (d * l) x r
        l, r = lb[i], lb[i + 1]
        q = _isometry(d * l, r, rng).reshape(d, l, r)
        arrs.append(q)
    arrs[-1] = arrs[-1] * np.asarray(sigma)[None, None, :]
    for j in range(N - h):             # right isometries: l x (d * r) rows orthonormal
        l, r = rb[j], rb[j + 1]
        q = _isometry(d * r, l, rng)   # (d r) x l, orthonormal columns
        arrs.append(q.T.reshape(l, d, r).transpose(1, 0, 2).copy())
    return arrs


def identity_mpo_arrays(N, d):
    """MPO arrays (left, right, physout, physin) of the identity, bond 1."""
    eye = np.eye(d).reshape(1, 1, d, d)
    return [eye.copy() for _ in range(N)]


def decaying_sigma(k, tau):
    """sigma_j = exp(-j / tau), j = 0..k-1."""
    return np.exp(-np.arange(k) / tau)
