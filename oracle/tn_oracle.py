"""CPU FP64 oracle for the contraction-plus-truncation path of tncontract.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product package ``tncontract_amd`` never imports it.

This is a NumPy restatement of the reference semantics, written for NumPy 2
(the reference itself calls ``np.product``/``np.float`` which NumPy 2 removed).
Each function cites the reference file:line whose behaviour it restates.  All
arithmetic is float64 through NumPy/LAPACK (``np.tensordot``, ``np.linalg.svd``
(gesdd), ``np.linalg.qr`` (geqrf+orgqr)), exactly the primitives the reference
calls.

How each function is pinned (checks live in ``tests/test_oracle.py``):
  * Tensor axis ops / contract / tensordot / trace / con: README and docstring
    known answers (shapes, labels), ``np.einsum`` brute force, closed form
    trace(A^N) for the ring network.
  * tensor_svd / truncated_svd / tensor_qr: reconstruction, orthogonality,
    singular values equal to the LAPACK values of the same reshaped matrix,
    docstring shape/label known answers, closed-form selection rule.
  * left_canonise / right_canonise / svd_compress: isometry checks, norm
    preservation, zero-state special case, and brute force against a
    sequential truncated SVD of the dense state vector (N <= 12).
  * contract_mps_mpo + XXZ MPO: dense Kronecker-sum Hamiltonian times the
    dense state; closed-form energy of the fully polarised state.
  * inner_product_mps / frob_distance_squared / expvals_mps: dense brute force,
    product-state closed forms.
  * tensor_lq: L.Q reconstruction, orthonormal Q rows, lower-triangular L.
  * ladder_contract (left / right / pairwise branches, intermediate
    contractions) and contract_virtual_indices: products of dense transfer
    matrices; contract_multi_index_tensor_with_one_dim_array through the
    square-lattice einsum brute force (tests/test_oracle_square_lattice.py).
  * variational_compress: exact at sufficient chi, never worse than its
    svd_compress start, left canonical, and the projection identities
    <w|v> = <w|w>, |v - w|^2 = |v|^2 - |w|^2 of the final single-site update.
  * tensor_to_matrix / matrix_to_tensor: element-by-element known answers
    with explicit index loops, and their round trip.
  * the keep rule of left_canonise's SVD path: two-site states with known
    Schmidt values scaled so that a relative and an absolute threshold keep
    different ranks; the result equals the dense Eckart-Young truncation.
  * init_mps_random: each site equals its np.random.rand draw divided by that
    draw's largest singular value (np.linalg.norm(., 2)).
  * reference fixture tests/golden/random_10site_mps.npz (extracted from the
    reference's tests/random_10site_mps.dat without unpickling): the
    invariants of tncontract/tests/test_canonical_form_mps_conversion.py.
Every function above has at least one of these checks: no function is
"parity unpinned".
"""
from __future__ import annotations

import copy as _copy
import uuid

import numpy as np

__all__ = [
    "OTensor", "contract", "tensordot", "tensor_product", "con", "distance",
    "tensor_to_matrix", "matrix_to_tensor", "tensor_svd", "truncated_svd",
    "tensor_qr", "OMPS", "OMPO", "contract_mps_mpo", "ladder_contract",
    "inner_product_mps", "frob_distance_squared", "expvals_mps",
    "svd_compress_mps", "left_canonical_form_mps", "right_canonical_form_mps",
    "unique_label", "init_mps_random",
]


def unique_label():
    """tncontract/label.py:89-92 -- a random uuid4 string."""
    return str(uuid.uuid4())


def _moveaxis_perm(nd, src, dst):
    """Axis permutation that moves axis ``src`` to position ``dst``."""
    order = list(range(nd))
    order.insert(dst, order.pop(src))
    return order


class OTensor:
    """tensor.py:16-591 -- ndarray plus one label per axis (first match wins)."""

    def __init__(self, data, labels=None, base_label="i"):
        self.data = np.array(data)
        if labels is None or len(labels) == 0:
            self.labels = [base_label + str(i) for i in range(self.data.ndim)]
        else:
            self.labels = labels

    # tensor.py:148-157
    @property
    def labels(self):
        return self._labels

    @labels.setter
    def labels(self, labels):
        if len(labels) != self.data.ndim:
            raise ValueError("Labels do not match shape of data.")
        self._labels = list(labels)

    @property
    def shape(self):
        return self.data.shape

    @property
    def rank(self):
        return self.data.ndim

    def __repr__(self):
        return "OTensor(shape=%r, labels=%r)" % (self.data.shape, self.labels)

    def __getitem__(self, *args):
        # tensor.py:142-145, 593-613
        return _ToContract(self, *args)

    def __mul__(self, other):
        out = self.copy()
        out.data = out.data * other
        return out

    __rmul__ = __mul__

    def __add__(self, other):
        # tensor.py:125-140
        a = self.copy()
        b = other.copy()
        a.consolidate_indices()
        b.consolidate_indices()
        return OTensor(a.data + b.data, labels=a.labels)

    def copy(self):
        # tensor.py:375-378
        return OTensor(self.data.copy(), _copy.copy(self.labels))

    def replace_label(self, old_labels, new_labels):
        # tensor.py:164-179
        old_labels = old_labels if isinstance(old_labels, list) else [old_labels]
        new_labels = new_labels if isinstance(new_labels, list) else [new_labels]
        for i, lab in enumerate(self.labels):
            if lab in old_labels:
                self.labels[i] = new_labels[old_labels.index(lab)]

    def move_index(self, label, position):
        # tensor.py:380-394: np.rollaxis in both branches is a single moveaxis
        src = self.labels.index(label)
        lab = self.labels.pop(src)
        self.labels.insert(position, lab)
        self.data = np.transpose(self.data, _moveaxis_perm(self.data.ndim, src, position))

    def move_indices(self, labels, position, preserve_relative_order=False):
        # tensor.py:396-482.  The moved axes are (a) in original axis order when
        # preserve_relative_order, else (b) grouped by label in the order of
        # `labels` (duplicates removed), each group in original axis order.
        labels = labels if isinstance(labels, list) else [labels]
        nd = self.data.ndim
        if preserve_relative_order:
            moved = [i for i in range(nd) if self.labels[i] in labels]
        else:
            seen = []
            for lab in labels:
                if lab not in seen:
                    seen.append(lab)
            moved = [i for lab in seen for i in range(nd) if self.labels[i] == lab]
        if position + len(moved) > nd:
            raise ValueError("Specified position too far right.")
        rest = [i for i in range(nd) if i not in moved]
        order = rest[:position] + moved + rest[position:]
        self.data = np.transpose(self.data, order)
        self.labels = [self.labels[i] for i in order]

    def consolidate_indices(self, labels=[]):
        # tensor.py:340-370: sorted unique labels (optionally filtered); the
        # axes carrying label number p are gathered at p, p+1, ... in their
        # current order and fused (first occurrence slowest).
        uniq = sorted(set(self.labels))
        if len(labels) != 0:
            uniq = [x for x in uniq if x in labels]
        for p, lab in enumerate(uniq):
            nd = self.data.ndim
            idx = [i for i in range(nd) if self.labels[i] == lab]
            rest = [i for i in range(nd) if i not in idx]
            order = rest[:p] + idx + rest[p:]
            data = np.transpose(self.data, order)
            shp = list(data.shape)
            fused = int(np.prod(shp[p:p + len(idx)]))
            self.data = np.reshape(data, shp[:p] + [fused] + shp[p + len(idx):])
            new_labels = [x for x in self.labels if x != lab]
            new_labels.insert(p, lab)
            self.labels = new_labels

    def contract_internal(self, label1, label2, index1=0, index2=0):
        # tensor.py:318-338
        a = [i for i, x in enumerate(self.labels) if x == label1][index1]
        b = [i for i, x in enumerate(self.labels) if x == label2][index2]
        self.data = np.trace(self.data, axis1=a, axis2=b)
        self.labels = [x for j, x in enumerate(self.labels) if j not in (a, b)]

    trace = contract_internal
    tr = contract_internal

    def add_dummy_index(self, label, position=0):
        # tensor.py:506-513
        self.data = self.data[np.newaxis]
        self.labels.insert(0, label)
        self.move_index(label, position)

    def remove_all_dummy_indices(self, labels=None):
        # tensor.py:515-529 (iterates over the ORIGINAL label list)
        orig_shape = self.shape
        for i, x in enumerate(list(self.labels)):
            if orig_shape[i] != 1:
                continue
            if labels is not None and x not in labels:
                continue
            self.move_index(x, 0)
            self.data = self.data[0]
            self.labels = self.labels[1:]

    def index_dimension(self, label):
        return self.data.shape[self.labels.index(label)]

    def conjugate(self):
        self.data = self.data.conjugate()

    def inv(self):
        # tensor.py:487-488
        self.data = np.linalg.inv(self.data)

    def prime_label(self, labels=None):
        # tensor.py:181-197 -- prime every label whose unprimed form is one of
        # `labels` (all labels when None); unprimed form = trailing "'" removed
        if labels is None:
            labels = list(self.labels)
        elif not isinstance(labels, list):
            labels = [labels]
        self.labels = [x + "'" if x.rstrip("'") in labels else x for x in self.labels]

    def unprime_label(self, labels=None):
        # tensor.py:199-226 -- remove one trailing prime from every label whose
        # unprimed form is one of `labels` (the tensor's labels when None)
        if labels is None:
            labels = list(self.labels)
        elif not isinstance(labels, list):
            labels = [labels]
        out = []
        for x in self.labels:
            if x.rstrip("'") in labels:
                if not x.endswith("'"):
                    raise ValueError("label is not primed")
                x = x[:-1]
            out.append(x)
        self.labels = out

    def norm(self):
        return np.linalg.norm(self.data)

    def contract(self, *args, **kwargs):
        t = contract(self, *args, **kwargs)
        self.data = t.data
        self.labels = t.labels

    def suf(self, suffix):
        t = self.copy()
        t.labels = [x + suffix for x in t.labels]
        return t


class _ToContract:
    def __init__(self, tensor, labels):
        self.tensor = tensor
        self.labels = labels

    def __mul__(self, other):
        l1 = list(self.labels) if isinstance(self.labels, tuple) else self.labels
        l2 = list(other.labels) if isinstance(other.labels, tuple) else other.labels
        return contract(self.tensor, other.tensor, l1, l2)


def _axes_for(labels_of_tensor, labels, index_slice):
    axes = []
    for lab in labels:
        axes.extend(i for i, x in enumerate(labels_of_tensor) if x == lab)
    if index_slice is not None:
        sl = [x if x != -1 else len(axes) - 1 for x in index_slice]
        axes = [a for i, a in enumerate(axes) if i in sl]
    return axes


def contract(tensor1, tensor2, labels1, labels2, index_slice1=None, index_slice2=None):
    """tensor.py:635-770 -- label-based tensordot; output = free labels of
    tensor1 then free labels of tensor2, each in original order."""
    labels1 = labels1 if isinstance(labels1, list) else [labels1]
    labels2 = labels2 if isinstance(labels2, list) else [labels2]
    ax1 = _axes_for(tensor1.labels, labels1, index_slice1)
    ax2 = _axes_for(tensor2.labels, labels2, index_slice2)
    if len(ax1) != len(ax2):
        raise ValueError("Number of indices in contraction does not match.")
    for i in range(len(ax1)):
        d1, d2 = tensor1.data.shape[ax1[i]], tensor2.data.shape[ax2[i]]
        if d1 != d2:
            raise ValueError(str(labels1[min(i, len(labels1) - 1)]) + " with dim=" + str(d1)
                             + " does not match " + str(labels2[min(i, len(labels2) - 1)])
                             + " with dim=" + str(d2))
    data = np.tensordot(tensor1.data, tensor2.data, (ax1, ax2))
    out_labels = ([x for j, x in enumerate(tensor1.labels) if j not in ax1]
                  + [x for j, x in enumerate(tensor2.labels) if j not in ax2])
    return OTensor(data, out_labels)


def tensordot(a, b, axes=2):
    """NumPy's tensordot (called by the reference at tensor.py:735)."""
    return np.tensordot(a, b, axes)


def tensor_product(*args):
    """tensor.py:773-779"""
    t = args[0]
    for x in args[1:]:
        t = contract(t, x, [], [])
    return t


def distance(tensor1, tensor2):
    """tensor.py:782-802"""
    t1, t2 = tensor1.copy(), tensor2.copy()
    t1.consolidate_indices()
    t2.consolidate_indices()
    if t1.labels != t2.labels:
        raise ValueError("Input tensors must have the same labels.")
    return np.linalg.norm(t1.data - t2.data)


def tensor_to_matrix(tensor, row_labels):
    """tensor.py:805-818"""
    t = tensor.copy()
    rows = 1
    for i, lab in enumerate(row_labels):
        t.move_index(lab, i)
        rows *= t.data.shape[i]
    return np.reshape(t.data, (rows, t.data.size // rows))


def matrix_to_tensor(matrix, shape, labels=None):
    """tensor.py:821-827"""
    return OTensor(np.reshape(np.array(matrix), shape), labels or [])


def _absorb(U, S, V, how):
    # tensor.py:945-956 (the reference hard-codes "svd_in"/"svd_out" here)
    if how == "left":
        return contract(U, S, ["svd_in"], ["svd_out"]), V
    if how == "right":
        return U, contract(S, V, ["svd_in"], ["svd_out"])
    sq = S.copy()
    sq.data = np.sqrt(sq.data)
    return contract(U, sq, ["svd_in"], ["svd_out"]), contract(sq, V, ["svd_in"], ["svd_out"])


def tensor_svd(tensor, row_labels, svd_label="svd_", absorb_singular_values=None):
    """tensor.py:830-959 -- rows moved first (in given order), thin SVD."""
    t = tensor.copy()
    rows = 1
    for i, lab in enumerate(row_labels):
        t.move_index(lab, i)
        rows *= t.data.shape[i]
    col_labels = [x for x in t.labels if x not in row_labels]
    old_shape = t.data.shape
    mat = np.reshape(t.data, (rows, t.data.size // rows))
    try:
        u, s, v = np.linalg.svd(mat, full_matrices=False)
    except (np.linalg.LinAlgError, ValueError):
        # tensor.py:916-929: retry with LAPACK gesvd; non-finite input raises
        # ValueError after printing what it holds
        import warnings
        import scipy.linalg
        warnings.warn('numpy.linalg.svd failed, trying scipy.linalg.svd with lapack_driver="gesvd"')
        try:
            u, s, v = scipy.linalg.svd(mat, full_matrices=False, lapack_driver="gesvd")
        except ValueError:
            print("tensor_svd failed. Matrix contains inf's: " + str(np.isinf(mat).any())
                  + ". Matrix contains nan's: " + str(np.isnan(mat).any()))
            raise
    U = OTensor(np.reshape(u, list(old_shape[:len(row_labels)]) + [u.shape[1]]),
                list(row_labels) + [svd_label + "in"])
    V = OTensor(np.reshape(v, [v.shape[0]] + list(old_shape[len(row_labels):])),
                [svd_label + "out"] + col_labels)
    S = OTensor(np.diag(s), [svd_label + "out", svd_label + "in"])
    if absorb_singular_values in ("left", "right", "both"):
        return _absorb(U, S, V, absorb_singular_values)
    return U, S, V


def truncated_svd(tensor, row_labels, chi=0, threshold=1e-15,
                  absorb_singular_values="right", absolute=True):
    """tensor.py:1112-1182 -- cap at chi first, then threshold (absolute, or
    relative to the largest value); truncated_evals = [thresholded, chi-cut]."""
    U, S, V = tensor_svd(tensor, row_labels)
    sv = np.diag(S.data)
    if chi:
        keep, cut_chi = sv[:chi], sv[chi:]
    else:
        keep, cut_chi = sv, np.array([])
    thr = threshold if absolute else sv[0] * threshold
    cut_thr = keep[keep <= thr]
    keep = keep[keep > thr]
    truncated_evals = np.concatenate((cut_thr, cut_chi), axis=0)
    S.data = np.diag(keep)
    U.move_index("svd_in", 0)
    U.data = U.data[:len(keep)]
    U.move_index("svd_in", len(U.labels) - 1)
    V.data = V.data[:len(keep)]
    if absorb_singular_values is None:
        return U, S, V
    how = absorb_singular_values if absorb_singular_values in ("left", "right") else "both"
    U2, V2 = _absorb(U, S, V, how)
    return U2, V2, truncated_evals


def tensor_qr(tensor, row_labels, qr_label="qr_"):
    """tensor.py:962-1055 -- move_indices(rows, 0), reduced Householder QR."""
    t = tensor.copy()
    row_labels = row_labels if isinstance(row_labels, list) else [row_labels]
    t.move_indices(row_labels, 0)
    rows = 1
    for i, lab in enumerate(t.labels):
        if lab not in row_labels:
            break
        rows *= t.data.shape[i]
    col_labels = [x for x in t.labels if x not in row_labels]
    old_shape = t.data.shape
    mat = np.reshape(t.data, (rows, t.data.size // rows))
    q, r = np.linalg.qr(mat, mode="reduced")
    Q = OTensor(np.reshape(q, list(old_shape[:len(row_labels)]) + [q.shape[1]]),
                row_labels + [qr_label + "in"])
    R = OTensor(np.reshape(r, [r.shape[0]] + list(old_shape[len(row_labels):])),
                [qr_label + "out"] + col_labels)
    return Q, R


def tensor_lq(tensor, row_labels, lq_label="lq_"):
    """tensor.py:1058-1109 -- a QR of the column labels, relabelled: returns
    (L, Q) with L labels row_labels + [lq_label + "in"], Q labels
    [lq_label + "out"] + the remaining labels."""
    col_labels = [x for x in tensor.labels if x not in row_labels]
    tmp = unique_label()
    Q, L = tensor_qr(tensor, col_labels, qr_label=tmp)
    Q.replace_label(tmp + "in", lq_label + "out")
    L.replace_label(tmp + "out", lq_label + "in")
    return L, Q


def con(*args):
    """tncon.py:4-159 -- internal traces first, then pairwise contractions in
    argument order (result stored at the lower tensor index), then the tensor
    product of the disconnected components."""
    tensors, pairs = [], []
    for x in args:
        if isinstance(x, list):
            (tensors if isinstance(x[0], OTensor) else pairs).extend(x)
        elif isinstance(x, OTensor):
            tensors.append(x)
        else:
            pairs.append(x)
    tensors = [t.copy() for t in tensors]
    flat = [lab for p in pairs for lab in p]
    if len(set(flat)) != len(flat):
        raise ValueError("Index found in more than one contraction pair.")
    owner = {}
    for i, t in enumerate(tensors):
        for lab in t.labels:
            if lab in owner:
                raise ValueError("Index label " + lab + " found in two tensors."
                                 + " Tensors must have unique index labelling.")
            owner[lab] = i
    internal, pairwise, keys = [], [], []
    for c in pairs:
        a, b = owner[c[0]], owner[c[1]]
        if a == b:
            internal.append(c)
            continue
        key = tuple(sorted((a, b)))
        if key in keys:
            k = keys.index((a, b))  # reference looks up the unsorted tuple
            if not isinstance(pairwise[k][0], list):
                pairwise[k][0] = [pairwise[k][0]]
                pairwise[k][1] = [pairwise[k][1]]
            pairwise[k][0].append(c[0])
            pairwise[k][1].append(c[1])
        else:
            pairwise.append(list(c))
            keys.append(key)
    for c in internal:
        tensors[owner[c[0]]].trace(c[0], c[1])
    comp = list(range(len(tensors)))
    for c in pairwise:
        first0 = c[0][0] if isinstance(c[0], list) else c[0]
        first1 = c[1][0] if isinstance(c[1], list) else c[1]
        d, e = owner[first0], owner[first1]
        if d == e:
            tensors[d].trace(c[0], c[1])
        else:
            if d < e:
                tensors[d] = contract(tensors[d], tensors[e], c[0], c[1])
                comp[e] = d
            else:
                tensors[e] = contract(tensors[e], tensors[d], c[1], c[0])
                comp[d] = e
            lo = min(d, e)
            for lab in tensors[lo].labels:
                owner[lab] = lo
    return tensor_product(*[tensors[comp.index(x)] for x in set(comp)])


# ---------------------------------------------------------------------------
# one-dimensional networks (onedim/onedim_core.py)
# ---------------------------------------------------------------------------
class _O1D:
    """onedim_core.py:31-183 -- object array of copied tensors; a dummy
    dimension-1 left/right index is inserted at axis 0 when missing."""

    def __init__(self, tensors, left_label="left", right_label="right"):
        self.left_label = left_label
        self.right_label = right_label
        self.data = [t.copy() for t in tensors]
        for x in self.data:
            if left_label not in x.labels:
                x.add_dummy_index(left_label)
            if right_label not in x.labels:
                x.add_dummy_index(right_label)

    def __len__(self):
        return len(self.data)

    def __iter__(self):
        return iter(self.data)

    def __getitem__(self, k):
        return self.data[k]

    def __setitem__(self, k, v):
        self.data[k] = v

    @property
    def nsites(self):
        return len(self.data)

    def reverse(self):
        self.data = self.data[::-1]
        self.left_label, self.right_label = self.right_label, self.left_label

    def complex_conjugate(self):
        for x in self.data:
            x.conjugate()

    def replace_labels(self, old, new):
        old = old if isinstance(old, list) else [old]
        new = new if isinstance(new, list) else [new]
        for x in self.data:
            x.replace_label(old, new)
        if self.left_label in old:
            self.left_label = new[old.index(self.left_label)]
        if self.right_label in old:
            self.right_label = new[old.index(self.right_label)]

    def unique_virtual_labels(self):
        self.replace_labels([self.left_label, self.right_label],
                            [unique_label(), unique_label()])

    def bonddims(self):
        if not self.data:
            return []
        b = [self.data[0].index_dimension(self.left_label)]
        b += [x.index_dimension(self.right_label) for x in self.data]
        return b

    def copy(self):
        return _O1D([x.copy() for x in self], self.left_label, self.right_label)


class OMPS(_O1D):
    """onedim_core.py:186-660 (MatrixProductState hot subset)."""

    def __init__(self, tensors, left_label="left", right_label="right", phys_label="phys"):
        super().__init__(tensors, left_label, right_label)
        self.phys_label = phys_label

    def copy(self):
        return OMPS([x.copy() for x in self], self.left_label, self.right_label, self.phys_label)

    def replace_labels(self, old, new):
        old = old if isinstance(old, list) else [old]
        new = new if isinstance(new, list) else [new]
        super().replace_labels(old, new)
        if self.phys_label in old:
            self.phys_label = new[old.index(self.phys_label)]

    def _zero_product_state(self):
        # onedim_core.py:274-278 / 305-309 / 323-330
        for k in range(len(self)):
            d = self[k].index_dimension(self.phys_label)
            self[k].data = np.zeros((d, 1, 1))
            self[k].labels = [self.phys_label, self.left_label, self.right_label]

    def left_canonise(self, start=0, end=-1, chi=None, threshold=1e-14,
                      normalise=False, qr_decomposition=False):
        """onedim_core.py:218-358"""
        N = len(self)
        if end == -1:
            end = N
        if qr_decomposition:
            for i in range(start, end):
                if i == N - 1:
                    nrm = np.linalg.norm(self[i].data)
                    if nrm == 0.0:
                        self._zero_product_state()
                        return
                    if normalise and start == 0:
                        self[i].data = self[i].data / nrm
                    return
                lab = unique_label()
                Q, R = tensor_qr(self[i], [self.phys_label, self.left_label], qr_label=lab)
                Q.replace_label(lab + "in", self.right_label)
                self[i] = Q
                self[i + 1] = contract(R, self[i + 1], self.right_label, self.left_label)
                self[i + 1].replace_label(lab + "out", self.left_label)
            return
        norm = 1.0
        for i in range(start, end):
            if i == N - 1:
                nrm = np.linalg.norm(self[i].data)
                if nrm == 0.0:
                    self._zero_product_state()
                    return
                if normalise and start == 0:
                    self[i].data = self[i].data / nrm
                else:
                    self[i].data = self[i].data * norm
                return
            lab = unique_label()
            U, S, V = tensor_svd(self[i], [self.phys_label, self.left_label], svd_label=lab)
            sv = np.diag(S.data)
            s0 = sv[0]
            if s0 == 0.0:
                self._zero_product_state()
                return
            sv = sv / s0
            norm *= s0
            keep = sv[sv > threshold]
            if chi:
                keep = keep[:chi]
            S.data = np.diag(keep)
            U.data = U.data[:, :, :len(keep)]
            V.data = V.data[:len(keep)]
            U.replace_label(lab + "in", self.right_label)
            self[i] = U
            nxt = contract(V, self[i + 1], self.right_label, self.left_label)
            nxt = contract(S, nxt, [lab + "in"], [lab + "out"])
            nxt.replace_label(lab + "out", self.left_label)
            self[i + 1] = nxt
            if i == end - 1:
                self[i + 1].data *= norm

    def right_canonise(self, start=0, end=-1, chi=None, threshold=1e-14,
                       normalise=False, qr_decomposition=False):
        """onedim_core.py:360-378"""
        self.reverse()
        N = len(self)
        if end == -1:
            end = N
        self.left_canonise(start=N - end, end=N - start, chi=chi, threshold=threshold,
                           normalise=normalise, qr_decomposition=qr_decomposition)
        self.reverse()

    def norm(self, canonical_form=False):
        """onedim_core.py:642-660"""
        if canonical_form == "left":
            return np.linalg.norm(self[-1].data)
        if canonical_form == "right":
            return np.linalg.norm(self[0].data)
        return np.sqrt(inner_product_mps(self, self))

    def svd_compress(self, chi=None, threshold=1e-15, normalise=False, reverse=False):
        """onedim_core.py:463-484"""
        if reverse:
            self.reverse()
        self.left_canonise(normalise=False, qr_decomposition=True)
        nrm = self.norm(canonical_form="left")
        self[-1].data /= nrm
        self.right_canonise(chi=chi, threshold=threshold, normalise=False)
        if not normalise:
            self[0].data *= nrm
        if reverse:
            self.reverse()

    def check_canonical_form(self, threshold=1e-14, print_output=False):
        """onedim_core.py:408-461 (return value only)."""
        cc = self.copy()
        cc.complex_conjugate()
        first_not_left = len(self) - 1
        for i in range(len(self) - 1):
            I = contract(self[i], cc[i], [self.phys_label, self.left_label],
                         [cc.phys_label, cc.left_label])
            if np.linalg.norm(I.data - np.identity(I.data.shape[0])) > threshold:
                first_not_left = i
                break
        first_not_right = 0
        for i in range(len(self) - 1, 0, -1):
            I = contract(self[i], cc[i], [self.phys_label, self.right_label],
                         [cc.phys_label, cc.right_label])
            if np.linalg.norm(I.data - np.identity(I.data.shape[0])) > threshold:
                first_not_right = i
                break
        return first_not_left, first_not_right


    def variational_compress(self, chi, max_iter=10, initial_guess=None, tolerance=1e-15,
                             normalise=False):
        """onedim_core.py:486-631 -- start from svd_compress(chi, reverse=True)
        (or the left-canonised initial guess), then alternate right-to-left
        and (mirrored) left-to-right single-site sweeps: each site is the
        contraction of the target with the left/right environments,
        right-canonised by an LQ decomposition whose L moves to the
        neighbour.  Converged when the relative changes of the norms along
        the second sweep are all below `tolerance`; returns a new MPS."""
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
                upd = contract(mps2[i], left_env[i - 1], mps2.left_label, le + "2")
                if i != mps2.nsites - 1:
                    upd = contract(upd, renv, mps2.right_label, re + "2")
                    upd.replace_label(re + "1", mps1.right_label)
                upd.replace_label([le + "1", mps2.phys_label], [mps1.left_label, mps1.phys_label])
                Lt, Q = tensor_lq(upd, mps1.left_label, lq_label=lq)
                Q.replace_label(lq + "out", mps1.left_label)
                Lt.replace_label(lq + "in", mps1.right_label)
                mps1[i] = Q
                mps1[i - 1] = contract(mps1[i - 1], Lt, mps1.right_label, mps1.left_label)
                norms.append(mps1[i - 1].norm())
                if i == mps2.nsites - 1:
                    renv = contract(conjugate(mps1[i]), mps2[i], mps1.phys_label,
                                    target.phys_label)
                    renv.remove_all_dummy_indices(labels=[mps1.right_label, mps2.right_label])
                else:
                    renv.contract(conjugate(mps1[i]), re + "1", mps1.right_label)
                    renv.contract(mps2[i], [mps1.phys_label, re + "2"],
                                  [target.phys_label, target.right_label])
                renv.replace_label([mps1.left_label, mps2.left_label], [re + "1", re + "2"])
                right_envs.append(renv.copy())
                if i == 1:
                    upd = contract(mps2[0], renv, mps2.right_label, re + "2")
                    upd.replace_label([mps2.phys_label, re + "1"],
                                      [mps1.phys_label, mps1.right_label])
                    mps1[0] = upd
            return right_envs, np.array(norms)

        for it in range(max_iter):
            left_env, norms1 = sweep(mps, self, left_env)
            mps.reverse()
            self.reverse()
            left_env, norms2 = sweep(mps, self, left_env)
            mps.reverse()
            self.reverse()
            if np.all(np.abs(norms2[1:] - norms2[:-1]) / norms2[1:] < tolerance):
                mps.replace_labels([mps.left_label, mps.right_label, mps.phys_label],
                                   [self.left_label, self.right_label, self.phys_label])
                if normalise:
                    mps[-1].data = mps[-1].data / mps.norm(canonical_form="left")
                return mps
            elif it == max_iter - 1:
                raise RuntimeError("variational_compress did not converge.")


class OMPO(_O1D):
    """onedim_core.py:1329-1367"""

    def __init__(self, tensors, left_label="left", right_label="right",
                 physout_label="physout", physin_label="physin"):
        super().__init__(tensors, left_label, right_label)
        self.physout_label = physout_label
        self.physin_label = physin_label


def contract_mps_mpo(mps, mpo):
    """onedim_core.py:1691-1708 -- per-site contraction over phys/physin, then
    consolidate_indices (alphabetical; MPS axis slower than MPO axis)."""
    if isinstance(mps, OMPSC):
        raise NotImplementedError("Function not implemented for" + "MatrixProductStateCanonical")
    out = []
    for i in range(len(mps)):
        t = contract(mps[i], mpo[i], mps.phys_label, mpo.physin_label)
        t.consolidate_indices()
        out.append(t)
    return OMPS(out, mps.left_label, mps.right_label, mpo.physout_label)


def ladder_contract(array1, array2, label1, label2, start=0, end=None,
                    complex_conjugate_array1=False, left_output_label="left",
                    right_output_label="right", return_intermediate_contractions=False):
    """onedim_core.py:1491-1663 -- all three branches: from the left when the
    interval starts at site 0, from the right when it ends at both arrays'
    last site, pairwise otherwise; optionally the list of partial
    contractions (in site order)."""
    if end is None:
        end = min(array1.nsites, array2.nsites) - 1
    if end < start:
        raise ValueError("Badly defined interval (end before start).")
    a1, a2 = array1.copy(), array2.copy()
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
                C = contract(a1[0], a2[0], rung, rung)
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
                C = contract(a1[end], a2[end], rung, rung)
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
            t = contract(a1[i], a2[i], rung, rung)
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


def inner_product_mps(mps_bra, mps_ket, complex_conjugate_bra=True, return_whole_tensor=False):
    """onedim_core.py:1666-1683 (canonical-form inputs go through their
    left-canonical form)"""
    bra = canonical_to_left_canonical(mps_bra) if isinstance(mps_bra, OMPSC) else mps_bra
    ket = canonical_to_left_canonical(mps_ket) if isinstance(mps_ket, OMPSC) else mps_ket
    t = ladder_contract(bra, ket, mps_bra.phys_label, mps_ket.phys_label,
                        complex_conjugate_array1=complex_conjugate_bra)
    return t if return_whole_tensor else t.data


def frob_distance_squared(mps1, mps2):
    """onedim_core.py:1686-1688"""
    ip = inner_product_mps
    return ip(mps1, mps1) + ip(mps2, mps2) - 2 * np.real(ip(mps1, mps2))


def contract_virtual_indices(array_1d, start=0, end=None, periodic_boundaries=True):
    """onedim_core.py:1399-1422 -- contract the virtual bonds of
    array_1d[start:end] (the slice excludes `end`), then, with periodic
    boundaries, trace the outer left/right indices."""
    C = array_1d[start].copy()
    for x in array_1d[start + 1:end]:
        C = contract(C, x, array_1d.right_label, array_1d.left_label)
    if periodic_boundaries:
        C.contract_internal(array_1d.right_label, array_1d.left_label)
    return C


def contract_multi_index_tensor_with_one_dim_array(tensor, array, label1, label2):
    """onedim_core.py:1370-1396 -- the label2 index of array[i] is contracted
    with the first still-open label1 index of `tensor` (the label1 indices are
    renamed to a temporary label for the duration), the virtual bonds along
    the way, and finally the array's outer left/right indices."""
    tmp = unique_label()
    tensor.replace_label(label1, tmp)
    C = contract(tensor, array[0], tmp, label2, index_slice1=[0])
    for i in range(1, len(array)):
        C = contract(C, array[i], [array.right_label, tmp], [array.left_label, label2],
                     index_slice1=[0, 1])
    C.contract_internal(array.right_label, array.left_label)
    tensor.replace_label(tmp, label1)
    return C


def left_canonical_form_mps(orig_mps, chi=0, threshold=1e-14, normalise=False):
    """onedim_core.py:1425-1436"""
    m = orig_mps.copy()
    m.left_canonise(chi=chi, threshold=threshold, normalise=normalise)
    return m


def right_canonical_form_mps(orig_mps, chi=0, threshold=1e-14, normalise=False):
    """onedim_core.py:1439-1445"""
    m = orig_mps.copy()
    m.right_canonise(chi=chi, threshold=threshold, normalise=normalise)
    return m


def svd_compress_mps(orig_mps, chi, threshold=1e-15, normalise=False):
    """onedim_core.py:1468-1473"""
    m = left_canonical_form_mps(orig_mps, threshold=threshold, normalise=normalise)
    return right_canonical_form_mps(m, chi=chi, threshold=threshold, normalise=normalise)


def variational_compress_mps(mps, chi, max_iter=10, initial_guess=None, tolerance=1e-15):
    """onedim_core.py:1474-1478"""
    return mps.variational_compress(chi, max_iter=max_iter, initial_guess=initial_guess,
                                    tolerance=tolerance)


def init_mps_random(nsites, physdim, bonddim=1, left_label="left", right_label="right",
                    phys_label="phys"):
    """onedim_utils.py:23-60: np.random.rand sites, each divided through its
    SVD over [phys, left] | right so that its largest singular value is 1."""
    if not np.iterable(physdim):
        physdim = [physdim] * nsites
    if not np.iterable(bonddim):
        bonddim = [bonddim] * (nsites - 1)
    bonddim = [1] + list(bonddim) + [1]
    tensors = []
    for i in range(nsites):
        rt = OTensor(np.random.rand(physdim[i], bonddim[i], bonddim[i + 1]),
                     [phys_label, left_label, right_label])
        U, S, V = tensor_svd(rt, [phys_label, left_label])
        S.data = S.data / S.data[0, 0]
        rt = contract(U, S, ["svd_in"], ["svd_out"])
        rt = contract(rt, V, ["svd_in"], ["svd_out"])
        tensors.append(rt)
    return OMPS(tensors, left_label=left_label, right_label=right_label, phys_label=phys_label)


def conjugate(tensor):
    """tensor.py:1185-1189"""
    t = tensor.copy()
    t.conjugate()
    return t


def expvals_mps(mps, oplist=[], sites=None, output_label=None, canonised=None):
    """onedim_utils.py:175-255"""
    if sites is None:
        sites = range(len(mps))
    if not np.iterable(sites):
        sites = [sites]
    n = len(sites)
    out = np.zeros(n, dtype=complex)
    ops = oplist if isinstance(oplist, list) else [oplist] * n
    if canonised == "left":
        mps.reverse()
        ops = ops[::-1]
    elif canonised != "right":
        mps.right_canonise()
    center = 0
    for i, site in enumerate(sites):
        mps.left_canonise(center, site)
        center = site
        op = ops[i]
        A = mps[site]
        if output_label is None:
            out_l, in_l = op.labels[0], op.labels[1]
        else:
            out_l = output_label
            in_l = [x for x in op.labels if x is not out_l][0]
        Ad = A.copy()
        Ad.conjugate()
        e = contract(A, op, mps.phys_label, in_l)
        e = contract(Ad, e, mps.phys_label, out_l)
        e.contract_internal(mps.left_label, mps.left_label, index1=0, index2=1)
        e.contract_internal(mps.right_label, mps.right_label, index1=0, index2=1)
        out[i] = e.data
    if canonised == "left":
        mps.reverse()
        out = out[::-1]
    return out


# ---------------------------------------------------------------- canonical form (Vidal)
class OMPSC(_O1D):
    """onedim_core.py:880-1326 -- Lambda Gamma Lambda ... Gamma Lambda: even
    positions hold the (diagonal) singular-value matrices, odd positions the
    Gamma tensors; physical site n sits at position 2n+1."""

    def __init__(self, tensors, left_label="left", right_label="right", phys_label="phys"):
        super().__init__(tensors, left_label, right_label)
        self.phys_label = phys_label

    def copy(self):
        return OMPSC([x.copy() for x in self], self.left_label, self.right_label, self.phys_label)

    def replace_labels(self, old, new):
        old = old if isinstance(old, list) else [old]
        new = new if isinstance(new, list) else [new]
        super().replace_labels(old, new)
        if self.phys_label in old:
            self.phys_label = new[old.index(self.phys_label)]

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

    @property
    def nsites_physical(self):
        return int((self.nsites - 1) / 2)

    def norm(self, canonical_form=True):
        """onedim_core.py:988-1002"""
        if canonical_form is True:
            return np.linalg.norm(self[-1].data) * np.linalg.norm(self[0].data)
        return np.sqrt(inner_product_mps(self, self))

    def check_canonical_form(self, threshold=1e-14, print_output=False):
        """onedim_core.py:1004-1065 -- Lambda.Gamma left isometries and
        Gamma.Lambda right isometries (edge sites may be unnormalised)."""
        not_left, not_right, not_norm = [], [], []
        n = self.nsites_physical
        for which, bad in (("left", not_left), ("right", not_right)):
            for i in range(n):
                p = self.physical_site(i)
                if which == "left":
                    A = self[p - 1][self.right_label, ] * self[p][self.left_label, ]
                    keep = [self.phys_label, self.left_label]
                else:
                    A = self[p][self.right_label, ] * self[p + 1][self.left_label, ]
                    keep = [self.phys_label, self.right_label]
                I = contract(A, conjugate(A), keep, keep)
                if np.linalg.norm(I.data - np.identity(I.data.shape[0])) > threshold:
                    if i == 0 or i == n - 1:
                        f = I.data.flatten()
                        if len(f[np.abs(f) > threshold]) > 1:
                            bad.append(i)
                        else:
                            not_norm.append(i)
                    else:
                        bad.append(i)
        return not_left, not_right, not_norm

    def _two_site_update(self, start, U, S, V):
        S1_inv = self[start].copy()
        S1_inv.inv()
        S2_inv = self[start + 4].copy()
        S2_inv.inv()
        self[start + 1] = S1_inv[self.right_label, ] * U[self.left_label, ]
        self[start + 2] = S
        self[start + 3] = V[self.right_label, ] * S2_inv[self.left_label, ]

    def compress_bond(self, singular_site, chi=None, threshold=1e-15):
        """onedim_core.py:1067-1095 (the early-exit test reads
        singulardim(singular_site(n)), i.e. position 4n, as the reference
        does)."""
        site = self.singular_site(singular_site)
        if self.singulardim(site) == 1:
            return
        start, end = site - 2, site + 2
        self[end - 1].prime_label(self.phys_label)
        t = contract_virtual_indices(self, start, end + 1, periodic_boundaries=False)
        S1_inv = self[start].copy()
        S1_inv.inv()
        S2_inv = self[end].copy()
        S2_inv.inv()
        U, S, V = truncated_svd(t, [self.phys_label, self.left_label], chi=chi,
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
        """onedim_core.py:1100-1186 -- contract Lambda Gamma (Lambda Gamma)
        Lambda, apply the gate, re-split by truncated SVD, divide the outer
        Lambdas back out."""
        gate_outputs, gate_inputs = self._gate_labels(gate, gate_outputs, gate_inputs)
        nsites = len(gate_inputs)
        if nsites > 2:
            raise NotImplementedError("gate acting on more than two sites.")
        start = self.physical_site(firstsite) - 1
        end = self.physical_site(firstsite + nsites - 1) + 1
        t = contract_virtual_indices(self, start, end + 1, periodic_boundaries=False)
        t = contract(t, gate, self.phys_label, gate_inputs)
        S1_inv = self[start].copy()
        S1_inv.inv()
        S2_inv = self[end].copy()
        S2_inv.inv()
        if nsites == 1:
            t.replace_label([gate_outputs[0]], [self.phys_label])
            t = S1_inv[self.right_label, ] * t[self.left_label, ]
            self[start + 1] = t[self.right_label, ] * S2_inv[self.left_label, ]
        elif nsites == 2:
            U, S, V = truncated_svd(t, [gate_outputs[0], self.left_label], chi=chi,
                                    threshold=threshold, absorb_singular_values=None)
            U.replace_label(["svd_in", gate_outputs[0]], [self.right_label, self.phys_label])
            V.replace_label(["svd_out", gate_outputs[1]], [self.left_label, self.phys_label])
            S.replace_label(["svd_out", "svd_in"], [self.left_label, self.right_label])
            self[start + 1] = S1_inv[self.right_label, ] * U[self.left_label, ]
            self[start + 2] = S
            self[start + 3] = V[self.right_label, ] * S2_inv[self.left_label, ]

    def swap_gate(self, i, chi=None, threshold=1e-15):
        """onedim_core.py:1188-1232 -- swap the physical indices of sites i and
        i+1 by re-splitting Lambda Gamma Lambda Gamma Lambda with the first
        site's physical index moved to the right factor."""
        start = self.physical_site(i) - 1
        end = self.physical_site(i + 1) + 1
        self[start + 1].prime_label(self.phys_label)
        t = contract_virtual_indices(self, start, end + 1, periodic_boundaries=False)
        S1_inv = self[start].copy()
        S1_inv.inv()
        S2_inv = self[end].copy()
        S2_inv.inv()
        U, S, V = truncated_svd(t, [self.left_label, self.phys_label], chi=chi,
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


def right_canonical_to_canonical(mps, threshold=1e-14):
    """onedim_core.py:1836-1884 -- left-to-right SVD sweep of a right-canonical
    MPS keeping singular values > threshold (absolute); Gamma = Lambda^-1 U."""
    N = mps.nsites
    S_prev = OTensor([[1.0]], labels=[mps.left_label, mps.right_label])
    S_prev_inv = S_prev.copy()
    B = mps[0]
    tensors = []
    lab = unique_label()
    for i in range(N):
        U, S, V = tensor_svd(B, [mps.phys_label, mps.left_label], svd_label=lab)
        sv = np.diag(S.data)
        keep = sv[sv > threshold]
        S.data = np.diag(keep)
        U.data = U.data[:, :, 0:len(keep)]
        V.data = V.data[0:len(keep)]
        U.replace_label(lab + "in", mps.right_label)
        V.replace_label(lab + "out", mps.left_label)
        S.replace_label([lab + "out", lab + "in"], [mps.left_label, mps.right_label])
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
    return OMPSC(tensors, mps.left_label, mps.right_label, mps.phys_label)


def left_canonical_to_canonical(mps, threshold=1e-14):
    """onedim_core.py:1887-1894"""
    r = mps.copy()
    r.reverse()
    c = right_canonical_to_canonical(r, threshold=threshold)
    c.reverse()
    return c


def canonical_to_right_canonical(mps):
    """onedim_core.py:1897-1911"""
    N = mps.nsites_physical
    tensors = []
    for i in range(N - 1):
        tensors.append(mps[mps.physical_site(i)][mps.right_label, ]
                       * mps[mps.physical_site(i) + 1][mps.left_label, ])
    tensors.append(mps[mps.physical_site(N - 1)])
    tensors[0].data = (tensors[0].data * np.linalg.norm(mps[-1].data)
                       * np.linalg.norm(mps[0].data))
    return OMPS(tensors, mps.left_label, mps.right_label, mps.phys_label)


def canonical_to_left_canonical(mps):
    """onedim_core.py:1914-1928"""
    N = mps.nsites_physical
    tensors = [mps[mps.physical_site(0)]]
    for i in range(1, N):
        tensors.append(mps[mps.physical_site(i) - 1][mps.right_label, ]
                       * mps[mps.physical_site(i)][mps.left_label, ])
    tensors[-1].data = (tensors[-1].data * np.linalg.norm(mps[-1].data)
                        * np.linalg.norm(mps[0].data))
    return OMPS(tensors, mps.left_label, mps.right_label, mps.phys_label)


def canonical_form_mps(orig_mps, chi=0, threshold=1e-14, normalise=False):
    """onedim_core.py:1448-1454"""
    mps = orig_mps.copy()
    mps.right_canonise(chi=chi, threshold=threshold, normalise=normalise)
    return right_canonical_to_canonical(mps, threshold=threshold)
