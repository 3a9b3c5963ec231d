"""Matrix product states/operators on the GPU (hot subset of
tncontract/onedim/onedim_core.py).

Same classes, methods, labels, axis orders and in-place mutation semantics
as the reference; every tensor is a device ``Tensor`` and every contraction,
QR, SVD, norm and scale runs in libtncontract_amd.  The sweeps stay in Python
(one C-ABI call per factorisation / contraction); the only host reads are the
singular values needed for truncation decisions and the scalar norms the
reference itself reads.
"""
from __future__ import annotations

import warnings

import numpy as np
import torch

from tncontract_amd import _lib as L
from tncontract_amd import tensor as tsr
from tncontract_amd.label import unique_label

__all__ = ["MatrixProductState", "MatrixProductOperator", "OneDimensionalTensorNetwork",
           "check_canonical_form_mps", "contract_mps_mpo", "frob_distance_squared",
           "inner_product_mps", "ladder_contract", "left_canonical_form_mps",
           "mps_complex_conjugate", "reverse_mps", "right_canonical_form_mps",
           "svd_compress_mps", "variational_compress_mps", "contract_virtual_indices",
           "contract_multi_index_tensor_with_one_dim_array", "MatrixProductStateCanonical",
           "right_canonical_to_canonical", "left_canonical_to_canonical",
           "canonical_to_right_canonical", "canonical_to_left_canonical", "canonical_form_mps"]


class OneDimensionalTensorNetwork:
    """onedim_core.py:31-183"""

    def __init__(self, tensors, left_label="left", right_label="right"):
        self.left_label = left_label
        self.right_label = right_label
        self.data = [x.copy() for x in tensors]
        for x in self.data:
            if left_label not in x.labels:
                x.add_dummy_index(left_label)
            if right_label not in x.labels:
                x.add_dummy_index(right_label)

    def __iter__(self):
        return iter(self.data)

    def __len__(self):
        return len(self.data)

    def __getitem__(self, key):
        return self.data[key]

    def __setitem__(self, key, value):
        self.data[key] = value

    def copy(self):
        return OneDimensionalTensorNetwork([x.copy() for x in self], self.left_label,
                                           self.right_label)

    def reverse(self):
        self.data = self.data[::-1]
        self.left_label, self.right_label = self.right_label, self.left_label

    def complex_conjugate(self):
        for x in self.data:
            x.conjugate()

    def replace_labels(self, old_labels, new_labels):
        if not isinstance(old_labels, list):
            old_labels = [old_labels]
        if not isinstance(new_labels, list):
            new_labels = [new_labels]
        for x in self.data:
            x.replace_label(old_labels, new_labels)
        if self.left_label in old_labels:
            self.left_label = new_labels[old_labels.index(self.left_label)]
        if self.right_label in old_labels:
            self.right_label = new_labels[old_labels.index(self.right_label)]

    def standard_virtual_labels(self, suffix=""):
        self.replace_labels([self.left_label, self.right_label],
                            ["left" + suffix, "right" + suffix])

    def unique_virtual_labels(self):
        self.replace_labels([self.left_label, self.right_label], [unique_label(), unique_label()])

    def leftdim(self, site):
        return self.data[site].index_dimension(self.left_label)

    def rightdim(self, site):
        return self.data[site].index_dimension(self.right_label)

    def bonddims(self):
        if self.nsites == 0:
            return []
        return [self.leftdim(0)] + [self.rightdim(i) for i in range(self.nsites)]

    @property
    def nsites(self):
        return len(self.data)

    @property
    def nsites_physical(self):
        return self.nsites


class MatrixProductState(OneDimensionalTensorNetwork):
    """onedim_core.py:186-660 (hot subset)."""

    def __init__(self, tensors, left_label="left", right_label="right", phys_label="phys"):
        OneDimensionalTensorNetwork.__init__(self, tensors, left_label, right_label)
        self.phys_label = phys_label

    def __repr__(self):
        return ("MatrixProductState(sites=%d, left_label=%r, right_label=%r, phys_label=%r)"
                % (len(self), self.left_label, self.right_label, self.phys_label))

    def copy(self):
        return MatrixProductState([x.copy() for x in self], self.left_label, self.right_label,
                                  self.phys_label)

    def replace_labels(self, old_labels, new_labels):
        if not isinstance(old_labels, list):
            old_labels = [old_labels]
        if not isinstance(new_labels, list):
            new_labels = [new_labels]
        OneDimensionalTensorNetwork.replace_labels(self, old_labels, new_labels)
        if self.phys_label in old_labels:
            self.phys_label = new_labels[old_labels.index(self.phys_label)]

    def standard_labels(self, suffix=""):
        self.replace_labels([self.left_label, self.right_label, self.phys_label],
                            ["left" + suffix, "right" + suffix, "phys" + suffix])

    def physdim(self, site):
        return self.data[site].index_dimension(self.phys_label)

    def physical_site(self, n):
        return n

    def _zero_product_state(self):
        # onedim_core.py:274-278 / 305-309 / 323-330
        for k in range(len(self)):
            d = self[k].index_dimension(self.phys_label)
            self[k] = tsr.Tensor(np.zeros((d, 1, 1)), [self.phys_label, self.left_label,
                                                      self.right_label])

    def left_canonise(self, start=0, end=-1, chi=None, threshold=1e-14, normalise=False,
                      qr_decomposition=False, _svd=None):
        """onedim_core.py:218-358.  `_svd` (internal): replaces the site SVD of
        the SVD path, called as _svd(site, row_labels, svd_label, i) and
        returning what tensor._tensor_svd returns, with at least as many
        singular values as the keep rule selects."""
        N = len(self)
        if end == -1:
            end = N
        if qr_decomposition:
            for i in range(start, end):
                if i == N - 1:
                    norm = self[i].norm()
                    if norm == 0.0:
                        self._zero_product_state()
                        return
                    if normalise and start == 0:
                        self[i].scale_(1.0 / norm)
                    return
                qr_label = unique_label()
                Q, R = tsr.tensor_qr(self[i], [self.phys_label, self.left_label], qr_label=qr_label)
                Q.replace_label(qr_label + "in", self.right_label)
                self[i] = Q
                self[i + 1] = tsr.contract(R, self[i + 1], self.right_label, self.left_label)
                self[i + 1].replace_label(qr_label + "out", self.left_label)
            return
        norm = 1.0
        for i in range(start, end):
            if i == N - 1:
                nrm = self[i].norm()
                if nrm == 0.0:
                    self._zero_product_state()
                    return
                if normalise and start == 0:
                    self[i].scale_(1.0 / nrm)
                else:
                    self[i].scale_(norm)
                return
            svd_label = unique_label()
            if _svd is None:
                U, S, V, s = tsr._tensor_svd(self[i], [self.phys_label, self.left_label],
                                             svd_label=svd_label)
            else:
                U, S, V, s = _svd(self[i], [self.phys_label, self.left_label], svd_label, i)
            sv = s.cpu().numpy()
            s0 = sv[0] if len(sv) else 0.0
            if s0 == 0.0:
                self._zero_product_state()
                return
            sv = sv / s0
            norm *= s0
            keep = sv[sv > threshold]
            if chi:
                keep = keep[:chi]
            k = len(keep)
            S = tsr.Tensor._wrap(L.diag_embed(_upload(keep), U.dtype), S.labels)
            U._data = U.data[:, :, :k]
            V._data = V.data[:k]
            U.replace_label(svd_label + "in", self.right_label)
            self[i] = U
            nxt = tsr.contract(V, self[i + 1], self.right_label, self.left_label)
            nxt = tsr.contract(S, nxt, [svd_label + "in"], [svd_label + "out"])
            nxt.replace_label(svd_label + "out", self.left_label)
            self[i + 1] = nxt
            if i == end - 1:
                self[i + 1].scale_(norm)

    def right_canonise(self, start=0, end=-1, chi=None, threshold=1e-14, normalise=False,
                       qr_decomposition=False, _svd=None):
        """onedim_core.py:360-378 (`_svd`: as in left_canonise, on the reversed
        chain)"""
        self.reverse()
        N = len(self)
        if end == -1:
            end = N
        self.left_canonise(start=N - end, end=N - start, chi=chi, threshold=threshold,
                           normalise=normalise, qr_decomposition=qr_decomposition, _svd=_svd)
        self.reverse()

    def norm(self, canonical_form=False):
        """onedim_core.py:642-660"""
        if canonical_form == "left":
            return self[-1].norm()
        if canonical_form == "right":
            return self[0].norm()
        return float(np.sqrt(inner_product_mps(self, self)))

    def svd_compress(self, chi=None, threshold=1e-15, normalise=False, reverse=False, _svd=None):
        """onedim_core.py:463-484: QR left sweep, normalise, truncating SVD
        right sweep, restore the norm.  `_svd` (internal): the right sweep's
        site SVD, as in right_canonise."""
        if reverse:
            self.reverse()
        self.left_canonise(normalise=False, qr_decomposition=True)
        norm = self.norm(canonical_form="left")
        if norm == 0.0:
            # the reference divides by the zero norm (NaN data, numpy's
            # "invalid value" warning); the right sweep's SVD then raises
            # ValueError for the NaN matrix (tensor.py:922-929)
            warnings.warn("invalid value encountered in divide", RuntimeWarning)
            self[-1].scale_(float("nan"))
        else:
            self[-1].scale_(1.0 / norm)
        self.right_canonise(chi=chi, threshold=threshold, normalise=False, _svd=_svd)
        if not normalise:
            self[0].scale_(norm)
        if reverse:
            self.reverse()

    def variational_compress(self, chi, max_iter=10, initial_guess=None, tolerance=1e-15,
                             normalise=False):
        """onedim_core.py:486-631: start from svd_compress(chi, reverse=True)
        (or the left-canonised `initial_guess`), then alternate right-to-left
        and mirrored single-site sweeps: each site becomes the contraction of
        the target with its left/right environments, right-canonised by an LQ
        decomposition whose L moves to the neighbour.  Converged when the
        relative changes of the site norms along the second sweep are all
        below `tolerance` (RuntimeError after `max_iter` full updates).
        Returns a new, left-canonical MPS with this MPS's labels."""
        mps, _, converged = _variational_compress(self, chi, max_iter, initial_guess, tolerance,
                                                  normalise)
        if not converged:
            raise RuntimeError("variational_compress did not converge.")
        return mps

    def check_canonical_form(self, threshold=1e-14, print_output=True):
        """onedim_core.py:408-461"""
        first_not_left = len(self) - 1
        for i in range(len(self) - 1):
            I = tsr.contract(self[i], self[i], [self.phys_label, self.left_label],
                             [self.phys_label, self.left_label])
            if _dist_identity(I) > threshold:
                first_not_left = i
                break
        first_not_right = 0
        for i in range(len(self) - 1, 0, -1):
            I = tsr.contract(self[i], self[i], [self.phys_label, self.right_label],
                             [self.phys_label, self.right_label])
            if _dist_identity(I) > threshold:
                first_not_right = i
                break
        if print_output:
            print("first site not left canonised: %d, first site not right canonised: %d"
                  % (first_not_left, first_not_right))
        return first_not_left, first_not_right


def _upload(arr):
    import torch
    return torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to(L.device())


def _dist_identity(I):
    """||I - 1||_F for a square device matrix (FP64 reductions on the GPU)."""
    n = I.shape[0]
    eye = L.diag_embed(_upload(np.ones(n)), I.dtype)
    return float(np.sqrt(L.sumsq(L.axpby(L.contiguous(I.data), eye, 1.0, -1.0))))


class MatrixProductStateCanonical(OneDimensionalTensorNetwork):
    """onedim_core.py:880-1326: Lambda Gamma Lambda ... Gamma Lambda -- even
    positions hold the diagonal singular-value matrices, odd positions the
    Gamma tensors (physical site n at position 2n+1).  Two-site updates
    contract Lambda Gamma Lambda Gamma Lambda on the GPU, re-split it by the
    device truncated SVD and divide the outer Lambdas back out (exact
    diagonal inverses on the device)."""

    def __init__(self, tensors, left_label="left", right_label="right", phys_label="phys"):
        OneDimensionalTensorNetwork.__init__(self, tensors, left_label=left_label,
                                             right_label=right_label)
        self.phys_label = phys_label

    def __repr__(self):
        return ("MatrixProductStateCanonical(tensors=%r, left_label=%r,right_label=%r, "
                "phys_label=%r)" % (self.data, self.left_label, self.right_label,
                                    self.phys_label))

    def __str__(self):
        return ("MatrixProductStateCanonical object: " + "sites (incl. singular value sites)= "
                + str(len(self)) + ", left_label = " + self.left_label + ", right_label = "
                + self.right_label + ", phys_label = " + self.phys_label)

    def copy(self):
        return MatrixProductStateCanonical([x.copy() for x in self], self.left_label,
                                           self.right_label, self.phys_label)

    def replace_labels(self, old_labels, new_labels):
        if not isinstance(old_labels, list):
            old_labels = [old_labels]
        if not isinstance(new_labels, list):
            new_labels = [new_labels]
        OneDimensionalTensorNetwork.replace_labels(self, old_labels, new_labels)
        if self.phys_label in old_labels:
            self.phys_label = new_labels[old_labels.index(self.phys_label)]

    def standard_labels(self, suffix=""):
        self.replace_labels([self.left_label, self.right_label, self.phys_label],
                            ["left" + suffix, "right" + suffix, "phys" + suffix])

    def physical_site(self, n):
        return 2 * n + 1

    def singular_site(self, n):
        return 2 * n

    def physdim(self, site):
        return self.data[self.physical_site(site)].index_dimension(self.phys_label)

    def singulardim(self, site):
        return self.data[self.singular_site(site)].index_dimension(self.left_label)

    def bonddims(self):
        return OneDimensionalTensorNetwork.bonddims(self)

    @property
    def nsites_physical(self):
        return int((self.nsites - 1) / 2)

    def norm(self, canonical_form=True):
        """onedim_core.py:988-1002"""
        if canonical_form is True:
            return self[-1].norm() * self[0].norm()
        return float(np.sqrt(inner_product_mps(self, self)))

    def check_canonical_form(self, threshold=1e-14, print_output=True):
        """onedim_core.py:1004-1065: Lambda.Gamma left isometries, Gamma.Lambda
        right isometries (identity residuals reduced on the GPU)."""
        not_left, not_right, not_normalised = [], [], []
        n = self.nsites_physical
        for side in ("left", "right"):
            bad = not_left if side == "left" else not_right
            for i in range(n):
                p = self.physical_site(i)
                if side == "left":
                    A = self[p - 1][self.right_label, ] * self[p][self.left_label, ]
                    keep = [self.phys_label, self.left_label]
                else:
                    A = self[p][self.right_label, ] * self[p + 1][self.left_label, ]
                    keep = [self.phys_label, self.right_label]
                I = tsr.contract(A, tsr.conjugate(A), keep, keep)
                if _dist_identity(I) > threshold:
                    if i == 0 or i == n - 1:
                        f = I.numpy().flatten()
                        if len(f[np.abs(f) > threshold]) > 1:
                            bad.append(i)
                        else:
                            not_normalised.append(i)
                    else:
                        bad.append(i)
        if print_output:
            if len(not_left) == 0 and len(not_right) == 0:
                if len(not_normalised) == 0:
                    print("MPS in canonical form (normalised)")
                else:
                    print("MPS in canonical form (unnormalised)")
            else:
                print("Physical sites not left-canonical:")
                print(not_left)
                print("Physical sites not right-canonical:")
                print(not_right)
        return not_left, not_right, not_normalised

    def _inverse(self, pos):
        t = self[pos].copy()
        t.inv()
        return t

    def _layer_inverses(self, positions):
        """{pos: self._inverse(pos)} with one host synchronisation: the
        diagonal inverses are enqueued back to back and their checks read
        together; a Lambda that is not diagonal, or is singular, goes through
        Tensor.inv (which raises LinAlgError for a singular one)."""
        pending, out = [], {}
        for pos in positions:
            d = self[pos].data
            if d.dim() == 2 and d.shape[0] == d.shape[1] and d.shape[0] > 0:
                inv_d, info = L.diag_inverse_async(d)
                pending.append((pos, inv_d, info))
            else:
                out[pos] = self._inverse(pos)
        if pending:
            infos = torch.stack([p[2] for p in pending]).cpu().numpy()
            for (pos, inv_d, _), (offdiag, zeros) in zip(pending, infos):
                if offdiag == 0 and zeros == 0:
                    out[pos] = tsr.Tensor._wrap(inv_d, list(self[pos].labels))
                else:
                    out[pos] = self._inverse(pos)
        return out

    def compress_bond(self, singular_site, chi=None, threshold=1e-15):
        """onedim_core.py:1067-1095.  As in the reference, the early-exit test
        reads singulardim(singular_site(n)), i.e. position 4n."""
        site = self.singular_site(singular_site)
        if self.singulardim(site) == 1:
            return
        start, end = site - 2, site + 2
        self[end - 1].prime_label(self.phys_label)
        t = contract_virtual_indices(self, start, end + 1, periodic_boundaries=False)
        inv = self._layer_inverses([start, end])
        S1_inv, S2_inv = inv[start], inv[end]
        U, S, V = tsr.truncated_svd(t, [self.phys_label, self.left_label], chi=chi,
                                    threshold=threshold, absorb_singular_values=None)
        U.replace_label("svd_in", self.right_label)
        V.replace_label("svd_out", self.left_label)
        S.replace_label(["svd_out", "svd_in"], [self.left_label, self.right_label])
        self[start + 1] = S1_inv[self.right_label, ] * U[self.left_label, ]
        self[start + 2] = S
        self[end - 1] = V[self.right_label, ] * S2_inv[self.left_label, ]
        self[end - 1].unprime_label(self.phys_label)

    def compress_all(self, chi=None, threshold=1e-15, normalise=False):
        raise NotImplementedError

    def _gate_labels(self, gate, gate_outputs, gate_inputs):
        if gate_outputs is None and gate_inputs is None:
            gate_outputs = gate.labels[:int(len(gate.labels) / 2)]
            gate_inputs = gate.labels[int(len(gate.labels) / 2):]
        elif gate_outputs is None:
            gate_outputs = [x for x in gate.labels if x not in gate_inputs]
        elif gate_inputs is None:
            gate_inputs = [x for x in gate.labels if x not in gate_outputs]
        if len(gate_outputs) != len(gate_inputs):
            raise ValueError("len(gate_outputs) != len(gate_inputs)")
        return gate_outputs, gate_inputs

    def apply_gate(self, gate, firstsite, gate_outputs=None, gate_inputs=None, chi=None,
                   threshold=1e-15):
        """onedim_core.py:1100-1186: 1- and 2-site gates, optimal truncation
        at the updated bond (exact SVD of the gated two-site block)."""
        gate_outputs, gate_inputs = self._gate_labels(gate, gate_outputs, gate_inputs)
        nsites = len(gate_inputs)
        if nsites > 2:
            raise NotImplementedError("gate acting on more than two sites.")
        start = self.physical_site(firstsite) - 1
        end = self.physical_site(firstsite + nsites - 1) + 1
        t = contract_virtual_indices(self, start, end + 1, periodic_boundaries=False)
        t = tsr.contract(t, gate, self.phys_label, gate_inputs)
        inv = self._layer_inverses([start, end])
        S1_inv, S2_inv = inv[start], inv[end]
        if nsites == 1:
            t.replace_label([gate_outputs[0]], [self.phys_label])
            t = S1_inv[self.right_label, ] * t[self.left_label, ]
            self[start + 1] = t[self.right_label, ] * S2_inv[self.left_label, ]
        elif nsites == 2:
            U, S, V = tsr.truncated_svd(t, [gate_outputs[0], self.left_label], chi=chi,
                                        threshold=threshold, absorb_singular_values=None)
            U.replace_label(["svd_in", gate_outputs[0]], [self.right_label, self.phys_label])
            V.replace_label(["svd_out", gate_outputs[1]], [self.left_label, self.phys_label])
            S.replace_label(["svd_out", "svd_in"], [self.left_label, self.right_label])
            self[start + 1] = S1_inv[self.right_label, ] * U[self.left_label, ]
            self[start + 2] = S
            self[start + 3] = V[self.right_label, ] * S2_inv[self.left_label, ]

    def apply_gate_layer(self, gate, firstsites, gate_outputs=None, gate_inputs=None, chi=None,
                         threshold=1e-15):
        """Extension (no counterpart in the reference): ``apply_gate(g, k,
        gate_outputs, gate_inputs, chi, threshold)`` for every two-site gate g
        (one Tensor for all bonds, or one per bond) and first site k in
        `firstsites`, on pairwise disjoint bonds (|k - k'| >= 2), with the
        bonds' SVDs in one batched launch.  Disjoint updates touch disjoint
        Gamma/Lambda positions and read only Lambdas no other update writes,
        so the result is that of the sequential calls."""
        from tncontract_amd.batched import batched_svd
        ks = list(firstsites)
        gates = list(gate) if isinstance(gate, (list, tuple)) else [gate] * len(ks)
        if len(gates) != len(ks):
            raise ValueError("one gate per bond, or one gate for all bonds")
        srt = sorted(ks)
        if any(b - a < 2 for a, b in zip(srt, srt[1:])):
            raise ValueError("apply_gate_layer: bonds must be disjoint (first sites >= 2 apart)")
        blocks = []
        for g, k in zip(gates, ks):
            go, gi = self._gate_labels(g, gate_outputs, gate_inputs)
            if len(gi) != 2:
                raise NotImplementedError("apply_gate_layer applies two-site gates")
            start = self.physical_site(k) - 1
            end = self.physical_site(k + 1) + 1
            t = contract_virtual_indices(self, start, end + 1, periodic_boundaries=False)
            t = tsr.contract(t, g, self.phys_label, gi)
            v, nrow = tsr._row_view(t, [go[0], self.left_label], move_all=False)
            shape = tuple(v.shape)
            m, n = int(np.prod(shape[:nrow])), int(np.prod(shape[nrow:]))
            blocks.append((k, go, list(v.labels), shape, nrow, L.materialize(v.data).view(m, n)))
        # the outer Lambdas' inverses (read-only in the layer, shared by
        # neighbouring bonds) before anything is written, as apply_gate takes
        # them before its writes
        inv = self._layer_inverses(sorted({self.physical_site(k) - 1 + o
                                           for k in ks for o in (0, 4)}))
        U, S, Vt, _ = batched_svd([b[5] for b in blocks])
        # every bond's singular values in one device-to-host copy, so the
        # updates below are enqueued without waiting on the device
        svals = torch.cat(S).cpu().numpy() if S else np.zeros(0)
        offs = np.cumsum([0] + [s.numel() for s in S])
        for (k, go, labels, shape, nrow, _), u, s, vt, o in zip(blocks, U, S, Vt, offs):
            sv = svals[o:o + s.numel()]
            kk = int(np.count_nonzero((sv[:chi] if chi else sv) > threshold))
            Ut = tsr.Tensor._wrap(L.contiguous(u[:, :kk]).view(shape[:nrow] + (kk,)),
                                  labels[:nrow] + ["svd_in"])
            Vm = tsr.Tensor._wrap(vt[:kk].view((kk,) + shape[nrow:]), ["svd_out"] + labels[nrow:])
            Sm = tsr.Tensor._wrap(L.diag_embed(s[:kk], u.dtype),
                                  [self.left_label, self.right_label])
            Ut.replace_label(["svd_in", go[0]], [self.right_label, self.phys_label])
            Vm.replace_label(["svd_out", go[1]], [self.left_label, self.phys_label])
            start = self.physical_site(k) - 1
            S1_inv = inv[start]
            S2_inv = inv[start + 4]
            self[start + 1] = S1_inv[self.right_label, ] * Ut[self.left_label, ]
            self[start + 2] = Sm
            self[start + 3] = Vm[self.right_label, ] * S2_inv[self.left_label, ]

    def swap_gate(self, i, chi=None, threshold=1e-15):
        """onedim_core.py:1188-1232: swap the physical indices of sites i and
        i+1 by re-splitting the two-site block with the first site's
        (primed) physical index moved to the right factor."""
        start = self.physical_site(i) - 1
        end = self.physical_site(i + 1) + 1
        self[start + 1].prime_label(self.phys_label)
        t = contract_virtual_indices(self, start, end + 1, periodic_boundaries=False)
        inv = self._layer_inverses([start, end])
        S1_inv, S2_inv = inv[start], inv[end]
        U, S, V = tsr.truncated_svd(t, [self.left_label, self.phys_label], chi=chi,
                                    threshold=threshold, absorb_singular_values=None)
        V.unprime_label(self.phys_label)
        U.replace_label("svd_in", self.right_label)
        V.replace_label("svd_out", self.left_label)
        S.replace_label(["svd_out", "svd_in"], [self.left_label, self.right_label])
        self[start + 1] = S1_inv[self.right_label, ] * U[self.left_label, ]
        self[start + 2] = S
        self[start + 3] = V[self.right_label, ] * S2_inv[self.left_label, ]

    def expval(self, gate, firstsite, gate_outputs=None, gate_inputs=None):
        """onedim_core.py:1234-1290"""
        gate_outputs, gate_inputs = self._gate_labels(gate, gate_outputs, gate_inputs)
        start = self.physical_site(firstsite) - 1
        end = self.physical_site(firstsite + len(gate_inputs) - 1) + 1
        t = contract_virtual_indices(self, start, end + 1, periodic_boundaries=False)
        td = t.copy()
        td.conjugate()
        exp = t[self.phys_label, ] * gate[gate_inputs]
        exp = td[self.phys_label, ] * exp[gate_outputs]
        exp.tr(self.left_label, self.left_label, index1=0, index2=1)
        exp.tr(self.right_label, self.right_label, index1=0, index2=1)
        return exp

    def ptrace(self, firstsite, lastsite=None):
        """onedim_core.py:1292-1326"""
        if lastsite is None:
            lastsite = firstsite
        start = self.physical_site(firstsite) - 1
        end = self.physical_site(lastsite) + 1
        t = contract_virtual_indices(self, start, end + 1, periodic_boundaries=False)
        td = t.copy()
        td.conjugate()
        for i, l in enumerate(t.labels):
            if l == self.phys_label:
                t.labels[i] = l + "_out" + str(i)
                td.labels[i] = l + "_in" + str(i)
        return t[self.left_label, self.right_label] * td[self.left_label, self.right_label]


class MatrixProductOperator(OneDimensionalTensorNetwork):
    """onedim_core.py:1329-1367"""

    def __init__(self, tensors, left_label="left", right_label="right", physout_label="physout",
                 physin_label="physin"):
        OneDimensionalTensorNetwork.__init__(self, tensors, left_label, right_label)
        self.physout_label = physout_label
        self.physin_label = physin_label

    def physoutdim(self, site):
        return self.data[site].index_dimension(self.physout_label)

    def physindim(self, site):
        return self.data[site].index_dimension(self.physin_label)


def _variational_compress(self, chi, max_iter, initial_guess, tolerance, normalise):
    """Body of MatrixProductState.variational_compress: returns the compressed
    MPS (relabelled, normalised on request), the number of full iterations
    run and whether the norm criterion was met."""
    if initial_guess is None:
        mps = self.copy()
        mps.svd_compress(chi=chi, reverse=True)
    else:
        mps = initial_guess
        mps.left_canonise(qr_decomposition=True)
    mps.replace_labels([mps.left_label, mps.right_label, mps.phys_label],
                       [unique_label(), unique_label(), unique_label()])
    le_label = unique_label()
    left_env = ladder_contract(mps, self, mps.phys_label, self.phys_label,
                               return_intermediate_contractions=True,
                               right_output_label=le_label, complex_conjugate_array1=True)
    target = self

    def sweep(mps1, mps2, left_env):
        le = left_env[0].labels[0][:-1]
        re = unique_label()
        lq = unique_label()
        right_envs = []
        norms = [mps1[-1].norm()]
        renv = None
        for i in range(mps2.nsites - 1, 0, -1):
            upd = tsr.contract(mps2[i], left_env[i - 1], mps2.left_label, le + "2")
            if i != mps2.nsites - 1:
                upd = tsr.contract(upd, renv, mps2.right_label, re + "2")
                upd.replace_label(re + "1", mps1.right_label)
            upd.replace_label([le + "1", mps2.phys_label], [mps1.left_label, mps1.phys_label])
            Lt, Q = tsr._lq_rowspace(upd, mps1.left_label, lq_label=lq)
            Q.replace_label(lq + "out", mps1.left_label)
            Lt.replace_label(lq + "in", mps1.right_label)
            mps1[i] = Q
            mps1[i - 1] = tsr.contract(mps1[i - 1], Lt, mps1.right_label, mps1.left_label)
            norms.append(mps1[i - 1].norm())
            if i == mps2.nsites - 1:
                renv = tsr.contract(tsr.conjugate(mps1[i]), mps2[i], mps1.phys_label,
                                    target.phys_label)
                renv.remove_all_dummy_indices(labels=[mps1.right_label, mps2.right_label])
            else:
                renv.contract(tsr.conjugate(mps1[i]), re + "1", mps1.right_label)
                renv.contract(mps2[i], [mps1.phys_label, re + "2"],
                              [target.phys_label, target.right_label])
            renv.replace_label([mps1.left_label, mps2.left_label], [re + "1", re + "2"])
            right_envs.append(renv.copy())
            if i == 1:
                upd = tsr.contract(mps2[0], renv, mps2.right_label, re + "2")
                upd.replace_label([mps2.phys_label, re + "1"],
                                  [mps1.phys_label, mps1.right_label])
                mps1[0] = upd
        return right_envs, np.array(norms)

    converged, it = False, -1
    for it in range(max_iter):
        left_env, norms1 = sweep(mps, self, left_env)
        mps.reverse()
        self.reverse()
        left_env, norms2 = sweep(mps, self, left_env)
        mps.reverse()
        self.reverse()
        if np.all(np.abs(norms2[1:] - norms2[:-1]) / norms2[1:] < tolerance):
            converged = True
            break
    mps.replace_labels([mps.left_label, mps.right_label, mps.phys_label],
                       [self.left_label, self.right_label, self.phys_label])
    if normalise:
        mps[-1].scale_(1.0 / mps.norm(canonical_form="left"))
    return mps, it + 1, converged


def contract_virtual_indices(array_1d, start=0, end=None, periodic_boundaries=True):
    """onedim_core.py:1399-1422"""
    C = array_1d[start].copy()
    for x in array_1d[start + 1:end]:
        C = tsr.contract(C, x, array_1d.right_label, array_1d.left_label)
    if periodic_boundaries:
        C.contract_internal(array_1d.right_label, array_1d.left_label)
    return C


def contract_multi_index_tensor_with_one_dim_array(tensor, array, label1, label2):
    """onedim_core.py:1370-1396: the label2 index of array[i] is contracted
    with the first still-open label1 index of `tensor` (temporarily renamed),
    together with the array's virtual bonds; finally the array's outer
    left/right indices are traced.  Each step is one MFMA contraction."""
    tmp = unique_label()
    tensor.replace_label(label1, tmp)
    C = tsr.contract(tensor, array[0], tmp, label2, index_slice1=[0])
    for i in range(1, len(array)):
        C = tsr.contract(C, array[i], [array.right_label, tmp], [array.left_label, label2],
                         index_slice1=[0, 1])
    C.contract_internal(array.right_label, array.left_label)
    tensor.replace_label(tmp, label1)
    return C


def left_canonical_form_mps(orig_mps, chi=0, threshold=1e-14, normalise=False):
    mps = orig_mps.copy()
    mps.left_canonise(chi=chi, threshold=threshold, normalise=normalise)
    return mps


def right_canonical_form_mps(orig_mps, chi=0, threshold=1e-14, normalise=False):
    mps = orig_mps.copy()
    mps.right_canonise(chi=chi, threshold=threshold, normalise=normalise)
    return mps


def reverse_mps(orig_mps):
    mps = orig_mps.copy()
    mps.reverse()
    return mps


def check_canonical_form_mps(mps, threshold=1e-14, print_output=True):
    return mps.check_canonical_form(threshold=threshold, print_output=print_output)


def svd_compress_mps(orig_mps, chi, threshold=1e-15, normalise=False):
    """onedim_core.py:1468-1473"""
    mps = left_canonical_form_mps(orig_mps, threshold=threshold, normalise=normalise)
    return right_canonical_form_mps(mps, chi=chi, threshold=threshold, normalise=normalise)


def variational_compress_mps(mps, chi, max_iter=10, initial_guess=None, tolerance=1e-15):
    """onedim_core.py:1474-1478"""
    return mps.variational_compress(chi, max_iter=max_iter, initial_guess=initial_guess,
                                    tolerance=tolerance)


def mps_complex_conjugate(mps):
    new = mps.copy()
    for x in new.data:
        x.conjugate()
    return new


def ladder_contract(array1, array2, label1, label2, start=0, end=None,
                    complex_conjugate_array1=False, left_output_label="left",
                    right_output_label="right", return_intermediate_contractions=False):
    """onedim_core.py:1491-1663: contract two 1D networks rung by rung (the
    label1/label2 index of every site pair) and along their virtual bonds.
    From the left when the interval starts at site 0, from the right when it
    ends at both arrays' last site, pairwise otherwise; optionally returns
    the list of partial contractions (in site order).  Every contraction is
    one MFMA GEMM with the label permutations fused (tn_contract)."""
    if end is None:
        end = min(array1.nsites, array2.nsites) - 1
    if end < start:
        raise ValueError("Badly defined interval (end before start).")
    a1 = OneDimensionalTensorNetwork(list(array1), array1.left_label, array1.right_label)
    a2 = OneDimensionalTensorNetwork(list(array2), array2.left_label, array2.right_label)
    if complex_conjugate_array1:
        a1.complex_conjugate()
    a1.unique_virtual_labels()
    a2.unique_virtual_labels()
    rung = unique_label()
    a1.replace_labels(label1, rung)
    a2.replace_labels(label2, rung)
    r1, r2 = right_output_label + "1", right_output_label + "2"
    l1, l2 = left_output_label + "1", left_output_label + "2"
    inter = []
    if start == 0:
        for i in range(0, end + 1):
            if i == 0:
                C = tsr.contract(a1[0], a2[0], rung, rung)
            else:
                C.contract(a1[i], a1.right_label, a1.left_label)
                C.contract(a2[i], [a2.right_label, rung], [a2.left_label, rung])
            if return_intermediate_contractions:
                t = C.copy()
                t.replace_label([a1.right_label, a2.right_label], [r1, r2])
                t.remove_all_dummy_indices(labels=[x for x in t.labels if x not in [r1, r2]])
                inter.append(t)
        C.replace_label([a1.right_label, a2.right_label], [r1, r2])
        C.remove_all_dummy_indices()
    elif end == a1.nsites - 1 and end == a2.nsites - 1:
        for i in range(end, start - 1, -1):
            if i == end:
                C = tsr.contract(a1[end], a2[end], rung, rung)
            else:
                C.contract(a1[i], a1.left_label, a1.right_label)
                C.contract(a2[i], [a2.left_label, rung], [a2.right_label, rung])
            if return_intermediate_contractions:
                t = C.copy()
                t.replace_label([a1.left_label, a2.left_label], [l1, l2])
                t.remove_all_dummy_indices(labels=[x for x in t.labels if x not in [l1, l2]])
                inter.insert(0, t)
        C.replace_label([a1.left_label, a2.left_label], [l1, l2])
        C.remove_all_dummy_indices()
    else:
        for i in range(start, end + 1):
            t = tsr.contract(a1[i], a2[i], rung, rung)
            if i == start:
                C = t
            else:
                C.contract(t, [a1.right_label, a2.right_label], [a1.left_label, a2.left_label])
            if return_intermediate_contractions:
                t = C.copy()
                t.replace_label([a1.right_label, a2.right_label, a1.left_label, a2.left_label],
                                [r1, r2, l1, l2])
                t.remove_all_dummy_indices(labels=[x for x in t.labels if x not in [r1, r2, l1, l2]])
                t.remove_all_dummy_indices()
                inter.append(t)
        C.replace_label([a1.right_label, a2.right_label, a1.left_label, a2.left_label],
                        [r1, r2, l1, l2])
        C.remove_all_dummy_indices()
    return inter if return_intermediate_contractions else C


def right_canonical_to_canonical(mps, threshold=1e-14):
    """onedim_core.py:1836-1884: left-to-right SVD sweep of a right-canonical
    MPS keeping singular values > threshold; Gamma_i = Lambda_i^-1 U_i."""
    N = mps.nsites
    S_prev = tsr.Tensor([[1.0]], labels=[mps.left_label, mps.right_label])
    S_prev_inv = S_prev.copy()
    B = mps[0]
    tensors = []
    svd_label = unique_label()
    for i in range(N):
        U, S, V, s = tsr._tensor_svd(B, [mps.phys_label, mps.left_label], svd_label=svd_label)
        sv = s.cpu().numpy()
        k = int(np.count_nonzero(sv > threshold))
        S = tsr.Tensor._wrap(L.diag_embed(_upload(sv[:k]), U.dtype), S.labels)
        U._data = U.data[:, :, 0:k]
        V._data = V.data[0:k]
        U.replace_label(svd_label + "in", mps.right_label)
        V.replace_label(svd_label + "out", mps.left_label)
        S.replace_label([svd_label + "out", svd_label + "in"], [mps.left_label, mps.right_label])
        G = S_prev_inv[mps.right_label, ] * U[mps.left_label, ]
        tensors.append(S_prev)
        tensors.append(G)
        if i == N - 1:
            tensors.append(S)
        else:
            V = S[mps.right_label, ] * V[mps.left_label, ]
            B = V[mps.right_label, ] * mps[i + 1][mps.left_label, ]
            S_prev = S.copy()
            S_prev_inv = S_prev.copy()
            S_prev_inv.inv()
    return MatrixProductStateCanonical(tensors, left_label=mps.left_label,
                                       right_label=mps.right_label, phys_label=mps.phys_label)


def left_canonical_to_canonical(mps, threshold=1e-14):
    """onedim_core.py:1887-1894"""
    mpsr = reverse_mps(mps)
    mpsc = right_canonical_to_canonical(mpsr, threshold=threshold)
    mpsc.reverse()
    return mpsc


def _rescale_shared(t, factor):
    """t's data replaced by factor * data (the reference rebinds .data on the
    tensor object, which may be shared with the source network)."""
    c = t.copy()
    c.scale_(factor)
    t._data = c._data


def canonical_to_right_canonical(mps):
    """onedim_core.py:1897-1911"""
    N = mps.nsites_physical
    tensors = []
    for i in range(N - 1):
        tensors.append(mps[mps.physical_site(i)][mps.right_label, ]
                       * mps[mps.physical_site(i) + 1][mps.left_label, ])
    tensors.append(mps[mps.physical_site(N - 1)])
    _rescale_shared(tensors[0], mps[-1].norm() * mps[0].norm())
    return MatrixProductState(tensors, left_label=mps.left_label, right_label=mps.right_label,
                              phys_label=mps.phys_label)


def canonical_to_left_canonical(mps):
    """onedim_core.py:1914-1928"""
    N = mps.nsites_physical
    tensors = [mps[mps.physical_site(0)]]
    for i in range(1, N):
        tensors.append(mps[mps.physical_site(i) - 1][mps.right_label, ]
                       * mps[mps.physical_site(i)][mps.left_label, ])
    _rescale_shared(tensors[-1], mps[-1].norm() * mps[0].norm())
    return MatrixProductState(tensors, left_label=mps.left_label, right_label=mps.right_label,
                              phys_label=mps.phys_label)


def canonical_form_mps(orig_mps, chi=0, threshold=1e-14, normalise=False):
    """onedim_core.py:1448-1454"""
    mps = orig_mps.copy()
    mps.right_canonise(chi=chi, threshold=threshold, normalise=normalise)
    return right_canonical_to_canonical(mps, threshold=threshold)


def inner_product_mps(mps_bra, mps_ket, complex_conjugate_bra=True, return_whole_tensor=False):
    """onedim_core.py:1666-1683 -- FP64 transfer-matrix sweep on the GPU
    (canonical-form inputs through their left-canonical form)."""
    bra = canonical_to_left_canonical(mps_bra) \
        if isinstance(mps_bra, MatrixProductStateCanonical) else mps_bra
    ket = canonical_to_left_canonical(mps_ket) \
        if isinstance(mps_ket, MatrixProductStateCanonical) else mps_ket
    t = ladder_contract(bra, ket, mps_bra.phys_label, mps_ket.phys_label,
                        complex_conjugate_array1=complex_conjugate_bra)
    if return_whole_tensor:
        return t
    return t.numpy()


def frob_distance_squared(mps1, mps2):
    """onedim_core.py:1686-1688"""
    ip = inner_product_mps
    return ip(mps1, mps1) + ip(mps2, mps2) - 2 * np.real(ip(mps1, mps2))


def contract_mps_mpo(mps, mpo):
    """onedim_core.py:1691-1708.  The consolidation permute (labels sorted,
    equal labels fused with the MPS axis slower) is fused into the store of
    the contraction kernel: each site is written once, already in its final
    (left, physout, right) layout."""
    if isinstance(mps, MatrixProductStateCanonical):
        raise NotImplementedError("Function not implemented for" + "MatrixProductStateCanonical")
    N = len(mps)
    new_mps = []
    for i in range(N):
        t = _contract_consolidated(mps[i], mpo[i], mps.phys_label, mpo.physin_label)
        new_mps.append(t)
    return MatrixProductState._from_owned(new_mps, mps.left_label, mps.right_label,
                                          mpo.physout_label)


def _contract_consolidated(t1, t2, l1, l2):
    """contract(t1, t2, l1, l2) followed by consolidate_indices(), computed as
    one kernel whose store offsets encode the consolidated layout."""
    import ctypes
    import torch
    ax1 = [i for i, x in enumerate(t1.labels) if x == l1]
    ax2 = [i for i, x in enumerate(t2.labels) if x == l2]
    free1 = [i for i in range(t1.rank) if i not in ax1]
    free2 = [i for i in range(t2.rank) if i not in ax2]
    labels = [t1.labels[i] for i in free1] + [t2.labels[i] for i in free2]
    dims = [t1.shape[i] for i in free1] + [t2.shape[i] for i in free2]
    # consolidated layout: simulate consolidate_indices on (labels, axis ids)
    groups = [[i] for i in range(len(labels))]
    labs = list(labels)
    for p, lab in enumerate(sorted(set(labels))):
        idx = [i for i in range(len(labs)) if labs[i] == lab]
        rest = [i for i in range(len(labs)) if i not in idx]
        fused = [a for i in idx for a in groups[i]]
        groups = [groups[i] for i in rest[:p]] + [fused] + [groups[i] for i in rest[p:]]
        labs = [labs[i] for i in rest[:p]] + [lab] + [labs[i] for i in rest[p:]]
    order = [a for g in groups for a in g]
    # stride of each uncontracted axis in the final dense layout
    stride = [0] * len(dims)
    acc = 1
    for a in reversed(order):
        stride[a] = acc
        acc *= dims[a]
    d1, d2 = tsr._common_dtype(t1.data, t2.data)
    out = torch.empty(acc, dtype=d1.dtype, device=d1.device)
    nf1 = len(free1)
    am = L.tensor_modes(d1, free1)
    ak = L.tensor_modes(d1, ax1)
    bk = L.tensor_modes(d2, ax2)
    bn = L.tensor_modes(d2, free2)
    cm = L.modes(dims[:nf1], stride[:nf1])
    cn = L.modes(dims[nf1:], stride[nf1:])
    if len(ax1) != len(ax2) or any(t1.shape[a] != t2.shape[b] for a, b in zip(ax1, ax2)):
        raise ValueError("contract_mps_mpo: physical dimensions do not match")
    L.contract(d1, am, ak, d2, bk, bn, out, cm, cn, L.dtype_code(d1), 0,
               int(np.prod([d1.shape[a] for a in free1])), int(np.prod([d2.shape[a] for a in free2])),
               int(np.prod([d1.shape[a] for a in ax1])))
    shape = [int(np.prod([dims[a] for a in g])) for g in groups]
    return tsr.Tensor._wrap(out.view(shape), labs)


def _from_owned(cls, tensors, left_label, right_label, phys_label):
    """Construct without the defensive per-site copy (tensors are fresh)."""
    obj = cls.__new__(cls)
    obj.left_label, obj.right_label, obj.phys_label = left_label, right_label, phys_label
    obj.data = list(tensors)
    for x in obj.data:
        if left_label not in x.labels:
            x.add_dummy_index(left_label)
        if right_label not in x.labels:
            x.add_dummy_index(right_label)
    return obj


MatrixProductState._from_owned = classmethod(_from_owned)
