"""Labelled device tensors and the contraction / decomposition API.

Mirrors the public surface of tncontract/tensor.py: ``Tensor`` (tensor.py:16-591),
``contract`` (:635-770), ``tensor_product`` (:773), ``distance`` (:782),
``tensor_to_matrix`` / ``matrix_to_tensor`` (:805-827), ``tensor_svd``
(:830-959), ``tensor_qr`` (:962-1055), ``truncated_svd`` (:1112-1182), plus a
NumPy-style ``tensordot``.

``Tensor.data`` is a torch tensor on the GPU (float64 or float32).  Axis moves
(``move_index``/``move_indices``) are metadata-only strided views; the
permutation is applied inside the next kernel that reads the tensor: the
contraction's tile loads, or the gather of the QR/SVD workspace.  All
arithmetic runs in libtncontract_amd (HIP, gfx950).
"""
from __future__ import annotations

import copy as _copy
import ctypes
import warnings

import numpy as np
import torch

from tncontract_amd import _lib as L
from tncontract_amd import label as lbl

__all__ = ["Tensor", "contract", "distance", "matrix_to_tensor", "tensor_to_matrix",
           "random_tensor", "tensor_product", "tensor_svd", "truncated_svd", "zeros_tensor",
           "tensordot"]


def _to_device(data):
    if isinstance(data, Tensor):
        data = data.data
    if isinstance(data, torch.Tensor):
        t = data
        if t.dtype not in (torch.float64, torch.float32):
            if t.is_complex():
                raise NotImplementedError("complex tensors are not supported on the GPU path")
            t = t.to(torch.float64)
        if t.device.type != "cuda":
            return t.to(L.device())
        return L.materialize(t)  # independent copy, like np.array(data)
    arr = np.asarray(data)
    if np.iscomplexobj(arr):
        raise NotImplementedError("complex tensors are not supported on the GPU path")
    if arr.dtype not in (np.float64, np.float32):
        arr = arr.astype(np.float64)
    return torch.from_numpy(np.ascontiguousarray(arr)).to(L.device())


class Tensor:
    """A device array plus one label per axis (tensor.py:16-591)."""

    def __init__(self, data, labels=None, base_label="i"):
        labels = [] if labels is None else labels
        self._data = _to_device(data)
        if len(labels) == 0:
            self.assign_labels(base_label=base_label)
        else:
            self.labels = labels

    @classmethod
    def _wrap(cls, t, labels):
        """Wrap an existing device tensor without copying it."""
        obj = cls.__new__(cls)
        obj._data = t
        obj.labels = labels
        return obj

    # ------------------------------------------------------------ data
    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, value):
        if isinstance(value, torch.Tensor) and value.device.type == "cuda" and \
                value.dtype in (torch.float64, torch.float32):
            self._data = value
        else:
            self._data = _to_device(value)

    def numpy(self):
        """Host copy of the data (numpy array, dense)."""
        return L.contiguous(self._data).cpu().numpy()

    @property
    def shape(self):
        return tuple(self._data.shape)

    @property
    def rank(self):
        return self._data.dim()

    @property
    def dtype(self):
        return self._data.dtype

    # ------------------------------------------------------------ labels
    def get_labels(self):
        return self._labels

    def set_labels(self, labels):
        if len(labels) == self._data.dim():
            self._labels = list(labels)
        else:
            raise ValueError("Labels do not match shape of data.")

    labels = property(get_labels, set_labels)

    def assign_labels(self, base_label="i"):
        self.labels = [base_label + str(i) for i in range(self._data.dim())]

    def replace_label(self, old_labels, new_labels):
        if not isinstance(old_labels, list):
            old_labels = [old_labels]
        if not isinstance(new_labels, list):
            new_labels = [new_labels]
        for i, label in enumerate(self.labels):
            if label in old_labels:
                self.labels[i] = new_labels[old_labels.index(label)]

    def prime_label(self, labels=None):
        if labels is None:
            labels = self.labels
        elif not isinstance(labels, list):
            labels = [labels]
        for i, label in enumerate(self.labels):
            for noprime in labels:
                if lbl.noprime_label(label) == noprime:
                    self.labels[i] = lbl.prime_label(self.labels[i])

    def unprime_label(self, labels=None):
        if labels is None:
            labels = self.labels
        elif not isinstance(labels, list):
            labels = [labels]
        for i, label in enumerate(self.labels):
            for noprime in labels:
                if lbl.noprime_label(label) == noprime:
                    self.labels[i] = lbl.unprime_label(self.labels[i])

    def add_suffix_to_labels(self, suffix):
        self.labels = [x + suffix for x in self.labels]

    def suf(self, suffix):
        t = self.copy()
        t.labels = [x + suffix for x in t.labels]
        return t

    # ------------------------------------------------------------ python protocol
    def __repr__(self):
        return "Tensor(shape=%r, labels=%r, dtype=%s)" % (self.shape, self.labels, self.dtype)

    def __str__(self):
        lines = ["Tensor object: ", "Data type: " + str(self.dtype),
                 "Number of indices: " + str(self.rank), "", "Index labels:"]
        for i, label in enumerate(self.labels):
            lines.append("   %d. (dim=%d) %s" % (i, self.shape[i], label))
        return "\n".join(lines)

    def __eq__(self, other):
        if not isinstance(other, Tensor):
            return False
        return (self.labels == other.labels and self.shape == other.shape
                and np.array_equal(self.numpy(), other.numpy()))

    def __mul__(self, other):
        # scalar multiple (tensor.py:95-108)
        try:
            alpha = float(other)
        except TypeError:
            raise TypeError("unsupported operand type(s) *: for '" + self.__class__.__name__
                            + "' and '" + other.__class__.__name__ + "'")
        out = self.copy()
        L.scale_(out._data, alpha)
        return out

    __rmul__ = __mul__

    def __add__(self, other):
        # tensor.py:125-140: consolidate both, then add
        try:
            a = self.copy()
            b = other.copy()
            a.consolidate_indices()
            b.consolidate_indices()
            if a.labels != b.labels or a.shape != b.shape:
                raise ValueError
            x, y = _common_dtype(a._data, b._data)
            return Tensor._wrap(L.axpby(x, y, 1.0, 1.0), a.labels)
        except Exception:
            raise TypeError("Can only add together tensors with the same"
                            + " indices: labels and dimensions of each index must match.")

    def __getitem__(self, *args):
        return ToContract(self, *args)

    # ------------------------------------------------------------ axis ops (views)
    def copy(self):
        """Independent dense copy (tensor.py:375-378)."""
        return Tensor._wrap(L.materialize(self._data), _copy.copy(self.labels))

    def move_index(self, label, position):
        """Move the first axis labelled `label` to `position` (tensor.py:380-394)."""
        index = self.labels.index(label)
        self.labels.pop(index)
        self.labels.insert(position, label)
        order = list(range(self._data.dim()))
        order.insert(position, order.pop(index))
        self._data = self._data.permute(order)

    def move_indices(self, labels, position, preserve_relative_order=False):
        """tensor.py:396-482 (same resulting order, one strided view)."""
        if not isinstance(labels, list):
            labels = [labels]
        nd = self._data.dim()
        if preserve_relative_order:
            moved = [i for i in range(nd) if self.labels[i] in labels]
        else:
            uniq = []
            for x in labels:
                if x not in uniq:
                    uniq.append(x)
            moved = [i for x in uniq for i in range(nd) if self.labels[i] == x]
        if position + len(moved) > nd:
            raise ValueError("Specified position too far right.")
        rest = [i for i in range(nd) if i not in moved]
        order = rest[:position] + moved + rest[position:]
        self._data = self._data.permute(order)
        self.labels = [self.labels[i] for i in order]

    def consolidate_indices(self, labels=[]):
        """Fuse equal labels; sort labels alphabetically (tensor.py:340-370).
        The label numbered p gathers its axes (in current order) at position
        p, exactly the sequence of moves the reference makes; the result is
        one strided view (a copy only if the fused axes are not viewable)."""
        uniq = sorted(set(self.labels))
        if len(labels) != 0:
            uniq = [x for x in uniq if x in labels]
        groups = [[i] for i in range(self._data.dim())]  # axis groups, in order
        labs = list(self.labels)
        for p, lab in enumerate(uniq):
            idx = [i for i in range(len(labs)) if labs[i] == lab]
            rest = [i for i in range(len(labs)) if i not in idx]
            fused = [a for i in idx for a in groups[i]]
            groups = [groups[i] for i in rest[:p]] + [fused] + [groups[i] for i in rest[p:]]
            labs = [labs[i] for i in rest[:p]] + [lab] + [labs[i] for i in rest[p:]]
        perm = [a for g in groups for a in g]
        shape = [int(np.prod([self._data.shape[a] for a in g])) for g in groups]
        self._data = _reshape(self._data.permute(perm), shape)
        self.labels = labs

    def sort_labels(self):
        self.consolidate_indices()

    def fuse_indices(self, indices_to_fuse, new_label, preserve_relative_order=False):
        """tensor.py:226-300"""
        self.move_indices(indices_to_fuse, 0, preserve_relative_order=preserve_relative_order)
        total, last = 1, 0
        for i, x in enumerate(self.labels):
            if x in indices_to_fuse:
                total *= self.shape[i]
                last = i
        self._data = _reshape(self._data, (total,) + self.shape[last + 1:])
        self.labels = [new_label] + self.labels[last + 1:]

    def split_index(self, label, new_dims, new_labels):
        """tensor.py:302-316"""
        if len(new_dims) != len(new_labels):
            raise ValueError("Length of new_dims must equal length of new_labels")
        i = self.labels.index(label)
        shape = self.shape[:i] + tuple(new_dims) + self.shape[i + 1:]
        self._data = _reshape(self._data, shape)
        self.labels = self.labels[:i] + list(new_labels) + self.labels[i + 1:]

    def add_dummy_index(self, label, position=0):
        """tensor.py:506-513"""
        self._data = self._data.unsqueeze(0)
        self.labels.insert(0, label)
        self.move_index(label, position)

    def remove_all_dummy_indices(self, labels=None):
        """tensor.py:515-529"""
        orig_shape = self.shape
        for i, x in enumerate(list(self.labels)):
            if orig_shape[i] != 1:
                continue
            if labels is not None and x not in labels:
                continue
            self.move_index(x, 0)
            self._data = self._data[0]
            self.labels = self.labels[1:]

    def index_dimension(self, label):
        return self.shape[self.labels.index(label)]

    def to_matrix(self, row_labels):
        return tensor_to_matrix(self, row_labels)

    # ------------------------------------------------------------ arithmetic
    def contract_internal(self, label1, label2, index1=0, index2=0):
        """Trace over two axes (tensor.py:318-338)."""
        a = [i for i, x in enumerate(self.labels) if x == label1][index1]
        b = [i for i, x in enumerate(self.labels) if x == label2][index2]
        t = self._data
        if t.shape[a] != t.shape[b]:
            raise ValueError("traced axes have different dimensions")
        keep = [i for i in range(t.dim()) if i not in (a, b)]
        out = torch.empty(tuple(t.shape[i] for i in keep), dtype=t.dtype, device=t.device)
        m = L.tensor_modes(t, keep)
        L.check(L.lib().tn_trace(L.ptr(t), ctypes.byref(m), t.shape[a], t.stride(a) + t.stride(b),
                                 L.ptr(out), L.dtype_code(t), L.stream_ptr()))
        self._data = out
        self.labels = [x for j, x in enumerate(self.labels) if j not in (a, b)]

    trace = contract_internal
    tr = contract_internal

    def contract(self, *args, **kwargs):
        t = contract(self, *args, **kwargs)
        self._data = t._data
        self.labels = t.labels

    def conjugate(self):
        """Real data: conjugation is the identity."""

    def norm(self):
        """Frobenius norm, FP64 reduction on the GPU (tensor.py:587-590)."""
        return float(np.sqrt(L.sumsq(self._data)))

    def scale_(self, alpha):
        """In place ``data *= alpha`` (materialises a strided view first)."""
        self._data = L.contiguous(self._data)
        L.scale_(self._data, alpha)
        return self

    def inv(self):
        """Matrix inverse of a 2-index tensor (tensor.py:487-488,
        np.linalg.inv).  A diagonal matrix (the singular-value matrices of
        MatrixProductStateCanonical) is inverted exactly on the device,
        1 / A_ii as LAPACK gives it; any other square matrix through the
        device SVD, V diag(1/s) U^T.  Singular input raises LinAlgError."""
        d = self._data
        if d.dim() != 2 or d.shape[0] != d.shape[1]:
            raise np.linalg.LinAlgError("Last 2 dimensions of the array must be square")
        n = d.shape[0]
        if n == 0:
            return
        out, info = L.diag_inverse(d)
        if info[0] == 0:
            if info[1] > 0:
                raise np.linalg.LinAlgError("Singular matrix")
            self._data = out
            return
        src = d if d.dtype == torch.float64 else L.convert(L.contiguous(d), torch.float64)
        u, s, vt = _svd_device(src, 1)
        w, winfo = L.diag_inverse(L.diag_embed(s))
        if winfo[1] > 0:
            raise np.linalg.LinAlgError("Singular matrix")
        inv = _contract_data(vt, [0], _contract_data(w, [1], u, [1]), [0])
        self._data = inv if d.dtype == torch.float64 else L.convert(inv, d.dtype)


class ToContract:
    """``A["a","b"] * B["c","d"]`` shorthand (tensor.py:593-613)."""

    def __init__(self, tensor, labels):
        self.tensor = tensor
        self.labels = labels

    def __mul__(self, other):
        l1 = list(self.labels) if isinstance(self.labels, tuple) else self.labels
        l2 = list(other.labels) if isinstance(other.labels, tuple) else other.labels
        return contract(self.tensor, other.tensor, l1, l2)


def _reshape(t, shape):
    """Reshape a strided view; copies through tn_permute only when the
    requested fusion is not expressible as a view."""
    shape = tuple(int(s) for s in shape)
    try:
        return t.view(shape)
    except RuntimeError:
        return L.materialize(t).view(shape)


def _common_dtype(a, b):
    if a.dtype == b.dtype:
        return a, b
    return L.convert(a, torch.float64), L.convert(b, torch.float64)


def random_tensor(*args, **kwargs):
    """tensor.py:616-621 (entries from numpy.random.rand, then uploaded)."""
    labels = kwargs.pop("labels", [])
    base_label = kwargs.pop("base_label", "i")
    return Tensor(np.random.rand(*args), labels=labels, base_label=base_label)


def zeros_tensor(*args, **kwargs):
    """tensor.py:624-632"""
    labels = kwargs.pop("labels", [])
    dtype = kwargs.pop("dtype", np.float64)
    base_label = kwargs.pop("base_label", "i")
    return Tensor(np.zeros(*args, dtype=dtype), labels=labels, base_label=base_label)


# ---------------------------------------------------------------- contraction
def _fuses(t, axes):
    """True when the axes, in this order and size-1 axes aside, fuse into one
    stride (each stride is the next one times the next extent)."""
    ax = [a for a in axes if t.shape[a] != 1]
    return all(t.stride(a) == t.stride(b) * t.shape[b] for a, b in zip(ax, ax[1:]))


def _contract_data(d1, ax1, d2, ax2, precision="native"):
    """tensordot of two device tensors over axis lists ax1/ax2: output axes are
    the free axes of d1 then those of d2 (each in order)."""
    if len(ax1) != len(ax2):
        raise ValueError("Number of indices in contraction does not match.")
    for a, b in zip(ax1, ax2):
        if d1.shape[a] != d2.shape[b]:
            raise ValueError("shape-mismatch for sum")
    d1, d2 = _common_dtype(d1, d2)
    free1 = [i for i in range(d1.dim()) if i not in ax1]
    free2 = [i for i in range(d2.dim()) if i not in ax2]
    out_shape = [d1.shape[i] for i in free1] + [d2.shape[i] for i in free2]
    dtype = d1.dtype
    prec = L.PREC[precision]
    if prec != 0 and dtype == torch.float64:
        d1, d2, dtype = L.convert(d1, torch.float32), L.convert(d2, torch.float32), torch.float32
    out = torch.empty(tuple(out_shape), dtype=dtype, device=d1.device)
    if out.numel() == 0:
        return out
    M = int(np.prod(out_shape[:len(free1)])) if free1 else 1
    N = int(np.prod(out_shape[len(free1):])) if free2 else 1
    K = int(np.prod([d1.shape[a] for a in ax1])) if ax1 else 1
    # An operand whose free or contracted axes do not fuse into one stride
    # makes the MFMA contraction decode its tile offsets through per-mode
    # tables (15-30 % slower on the sweep's products).  For a GEMM-sized
    # product (every extent >= 64, >= 512 flops per operand element) one
    # permuted dense copy in (free, contracted) / (contracted, free) order
    # costs less, as the transpose inside np.tensordot (tensor.py:735) does.
    if min(M, N, K) >= 64:
        flops = 2.0 * M * N * K
        if (not (_fuses(d1, free1) and _fuses(d1, ax1)) and d1.numel() >= 65536
                and flops >= 512.0 * d1.numel()):
            d1 = L.contiguous(d1.permute(*(free1 + ax1)))
            free1, ax1 = list(range(len(free1))), list(range(len(free1), d1.dim()))
        if (not (_fuses(d2, ax2) and _fuses(d2, free2)) and d2.numel() >= 65536
                and flops >= 512.0 * d2.numel()):
            d2 = L.contiguous(d2.permute(*(ax2 + free2)))
            ax2, free2 = list(range(len(ax2))), list(range(len(ax2), d2.dim()))
    # dense output strides
    strides, acc = [0] * len(out_shape), 1
    for i in range(len(out_shape) - 1, -1, -1):
        strides[i] = acc
        acc *= out_shape[i]
    nf1 = len(free1)
    am = L.tensor_modes(d1, free1)
    ak = L.tensor_modes(d1, ax1)
    bk = L.tensor_modes(d2, ax2)
    bn = L.tensor_modes(d2, free2)
    cm = L.modes(out_shape[:nf1], strides[:nf1])
    cn = L.modes(out_shape[nf1:], strides[nf1:])
    L.contract(d1, am, ak, d2, bk, bn, out, cm, cn, L.dtype_code(d1), prec, M, N, K)
    return out


def tensordot(a, b, axes=2):
    """NumPy ``tensordot`` semantics on device tensors (the primitive the
    reference calls at tensor.py:735).  ``a``/``b``: Tensor, torch or numpy."""
    da = a.data if isinstance(a, Tensor) else _to_device(a)
    db = b.data if isinstance(b, Tensor) else _to_device(b)
    if isinstance(axes, int):
        ax1 = list(range(da.dim() - axes, da.dim()))
        ax2 = list(range(axes))
    else:
        ax1, ax2 = axes
        ax1 = [ax1] if isinstance(ax1, int) else list(ax1)
        ax2 = [ax2] if isinstance(ax2, int) else list(ax2)
    ax1 = [x % da.dim() for x in ax1]
    ax2 = [x % db.dim() for x in ax2]
    return _contract_data(da, ax1, db, ax2)


def contract(tensor1, tensor2, labels1, labels2, index_slice1=None, index_slice2=None,
             precision="native"):
    """tensor.py:635-770.  ``precision`` ("native" | "fp32" | "bf16x3") is an
    extension: the arithmetic of the MFMA contraction."""
    if not isinstance(labels1, list):
        labels1 = [labels1]
    if not isinstance(labels2, list):
        labels2 = [labels2]
    ax1 = []
    for label in labels1:
        ax1.extend(i for i, x in enumerate(tensor1.labels) if x == label)
    ax2 = []
    for label in labels2:
        ax2.extend(i for i, x in enumerate(tensor2.labels) if x == label)
    if index_slice1 is not None:
        sl = [x if x != -1 else len(ax1) - 1 for x in index_slice1]
        ax1 = [j for i, j in enumerate(ax1) if i in sl]
    if index_slice2 is not None:
        sl = [x if x != -1 else len(ax2) - 1 for x in index_slice2]
        ax2 = [j for i, j in enumerate(ax2) if i in sl]
    if len(ax1) != len(ax2):
        raise ValueError("Number of indices in contraction does not match.")
    for i in range(len(ax1)):
        d1 = tensor1.shape[ax1[i]]
        d2 = tensor2.shape[ax2[i]]
        if d1 != d2:
            raise ValueError(str(labels1[min(i, len(labels1) - 1)]) + " with dim=" + str(d1)
                             + " does not match " + str(labels2[min(i, len(labels2) - 1)])
                             + " with dim=" + str(d2))
    out = _contract_data(tensor1.data, ax1, tensor2.data, ax2, precision)
    labels = ([x for j, x in enumerate(tensor1.labels) if j not in ax1]
              + [x for j, x in enumerate(tensor2.labels) if j not in ax2])
    return Tensor._wrap(out, labels)


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
    x, y = _common_dtype(t1.data, t2.data)
    return float(np.sqrt(L.sumsq(L.axpby(x, y, 1.0, -1.0))))


def tensor_to_matrix(tensor, row_labels):
    """tensor.py:805-818 (returns a dense device matrix)."""
    t = Tensor._wrap(tensor.data, list(tensor.labels))
    rows = 1
    for i, label in enumerate(row_labels):
        t.move_index(label, i)
        rows *= t.shape[i]
    cols = t.data.numel() // max(rows, 1)
    return L.materialize(t.data).view(rows, cols)


def matrix_to_tensor(matrix, shape, labels=None):
    """tensor.py:821-827"""
    labels = [] if labels is None else labels
    t = Tensor(matrix)
    t._data = _reshape(t._data, shape)
    t.labels = labels if labels else ["i" + str(i) for i in range(len(shape))]
    return t


# ---------------------------------------------------------------- decompositions
def _row_view(tensor, row_labels, move_all):
    """A view of `tensor` with the row labels first.  Returns (view, nrows)."""
    t = Tensor._wrap(tensor.data, list(tensor.labels))
    if move_all:
        t.move_indices(row_labels, 0)
        nrow = 0
        for lab in t.labels:
            if lab not in row_labels:
                break
            nrow += 1
    else:
        for i, label in enumerate(row_labels):
            t.move_index(label, i)
        nrow = len(row_labels)
    return t, nrow


def _matrix_modes(d, nrow):
    rows = L.tensor_modes(d, range(nrow))
    cols = L.tensor_modes(d, range(nrow, d.dim()))
    m = int(np.prod(d.shape[:nrow])) if nrow else 1
    n = int(np.prod(d.shape[nrow:])) if d.dim() > nrow else 1
    return rows, cols, m, n


# Jacobi sweep limits of the SVD: the library's default for the first
# attempt, and the retry that stands in for the reference's second LAPACK
# driver (tensor.py:914-921) when the first attempt does not converge.
SVD_MAX_SWEEPS = 40
SVD_RETRY_SWEEPS = 160


def _svd_nonfinite(d):
    """tensor.py:922-929: report which non-finite entries the matrix holds and
    raise ValueError (what scipy's gesvd raises for them)."""
    ninf, nnan = L.count_nonfinite(d)
    print("tensor_svd failed. Matrix contains inf's: " + str(ninf > 0)
          + ". Matrix contains nan's: " + str(nnan > 0))
    raise ValueError("array must not contain infs or NaNs")


def _svd_device(d, nrow, max_sweeps=0):
    """Thin SVD of the (rows | cols) matrix view of d.  Returns (u, s, vt) with
    u: m x k, s: k (float64, descending), vt: k x n, all on the device.

    Failure handling follows tensor.py:914-929: a Jacobi that does not
    converge in `max_sweeps` (0: SVD_MAX_SWEEPS) is retried once with
    SVD_RETRY_SWEEPS after a warning, LinAlgError if that fails too; a matrix
    with infinite or NaN entries prints the reference's diagnostic and raises
    ValueError."""
    rows, cols, m, n = _matrix_modes(d, nrow)
    k = min(m, n)
    dt = d.dtype
    u = torch.empty((m, k), dtype=dt, device=d.device)
    s = torch.empty((k,), dtype=torch.float64, device=d.device)
    vt = torch.empty((k, n), dtype=dt, device=d.device)
    if k == 0:
        return u, s, vt
    lib = L.lib()
    ws = L.workspace(lib.tn_svd_workspace(m, n, L.dtype_code(d)))
    sweeps = ctypes.c_int(0)

    def run(limit):
        return lib.tn_svd(L.ptr(d), ctypes.byref(rows), ctypes.byref(cols), m, n, L.ptr(u),
                          L.ptr(s), L.ptr(vt), L.dtype_code(d), 0.0, int(limit),
                          ctypes.byref(sweeps), L.ptr(ws), ws.numel(), L.stream_ptr())

    limit = max_sweeps if max_sweeps > 0 else SVD_MAX_SWEEPS
    rc = run(limit)
    if rc == L.TN_ERR_NOCONV:
        if any(L.count_nonfinite(d)):
            _svd_nonfinite(d)
        warnings.warn("tn_svd: the Jacobi SVD did not converge in %d sweeps, trying again with %d"
                      % (limit, SVD_RETRY_SWEEPS))
        rc = run(SVD_RETRY_SWEEPS)
    L.check(rc)
    if not np.all(np.isfinite(s.cpu().numpy())):
        _svd_nonfinite(d)
    return u, s, vt


def _eigh_psd_device(d, nrow, tol=0.0, info=None):
    """Eigendecomposition of the symmetric positive semi-definite (rows | cols)
    matrix view of d: returns (u, lam) with u: n x n float64 (column j =
    eigenvector j) and lam: n float64, descending.  tol: Jacobi orthogonality
    tolerance (0: n * eps); info (dict, optional) receives "sweeps"."""
    rows, cols, m, n = _matrix_modes(d, nrow)
    if m != n:
        raise ValueError("eigh_psd needs a square matrix, got %d x %d" % (m, n))
    u = torch.empty((n, n), dtype=torch.float64, device=d.device)
    lam = torch.empty((n,), dtype=torch.float64, device=d.device)
    if n == 0:
        return u, lam
    lib = L.lib()
    ws = L.workspace(lib.tn_eigh_psd_workspace(n))
    sweeps = ctypes.c_int(0)
    rc = lib.tn_eigh_psd(L.ptr(d), ctypes.byref(rows), ctypes.byref(cols), n, L.ptr(u),
                         L.ptr(lam), L.dtype_code(d), float(tol), 0, ctypes.byref(sweeps),
                         L.ptr(ws), ws.numel(), L.stream_ptr())
    L.check(rc)
    if info is not None:
        info["sweeps"] = sweeps.value
    return u, lam


def _absorb(U, S, V, how):
    # tensor.py:945-956 (the reference hard-codes "svd_in"/"svd_out" here)
    if how == "left":
        return contract(U, S, ["svd_in"], ["svd_out"]), V
    if how == "right":
        return U, contract(S, V, ["svd_in"], ["svd_out"])
    sq = S.copy()
    sq._data = _sqrt_diag(sq._data)
    return contract(U, sq, ["svd_in"], ["svd_out"]), contract(sq, V, ["svd_in"], ["svd_out"])


def _sqrt_diag(Sdata):
    d = torch.diagonal(Sdata)
    # a fresh copy (materialize / convert): the square root is taken in place
    s = L.sqrt_(L.materialize(d) if d.dtype == torch.float64 else L.convert(d, torch.float64))
    return L.diag_embed(s, Sdata.dtype)


def _tensor_svd(tensor, row_labels, svd_label="svd_"):
    t, nrow = _row_view(tensor, row_labels, move_all=False)
    col_labels = [x for x in t.labels if x not in row_labels]
    old_shape = t.shape
    u, s, vt = _svd_device(t.data, nrow)
    k = s.numel()
    U = Tensor._wrap(u.view(tuple(old_shape[:nrow]) + (k,)), list(row_labels) + [svd_label + "in"])
    V = Tensor._wrap(vt.view((k,) + tuple(old_shape[nrow:])), [svd_label + "out"] + col_labels)
    S = Tensor._wrap(L.diag_embed(s, u.dtype), [svd_label + "out", svd_label + "in"])
    return U, S, V, s


def tensor_svd(tensor, row_labels, svd_label="svd_", absorb_singular_values=None):
    """tensor.py:830-959 -- one-sided block Jacobi SVD on the GPU."""
    U, S, V, _ = _tensor_svd(tensor, row_labels, svd_label)
    if absorb_singular_values in ("left", "right", "both"):
        return _absorb(U, S, V, absorb_singular_values)
    return U, S, V


def truncated_svd(tensor, row_labels, chi=0, threshold=1e-15, absorb_singular_values="right",
                  absolute=True):
    """tensor.py:1112-1182: cap at chi, then threshold (absolute, or relative
    to the largest singular value); truncated_evals = [thresholded, chi-cut]."""
    U, S, V, s_dev = _tensor_svd(tensor, row_labels)
    sv = s_dev.cpu().numpy()
    if chi:
        keep, cut_chi = sv[:chi], sv[chi:]
    else:
        keep, cut_chi = sv, np.array([])
    thr = threshold if absolute else sv[0] * threshold
    cut_thr = keep[keep <= thr]
    keep = keep[keep > thr]
    truncated_evals = np.concatenate((cut_thr, cut_chi), axis=0)
    k = len(keep)
    S._data = L.contiguous(S.data[:k, :k])
    U.move_index("svd_in", 0)
    U._data = U.data[:k]
    U.move_index("svd_in", len(U.labels) - 1)
    V._data = V.data[:k]
    if absorb_singular_values is None:
        return U, S, V
    how = absorb_singular_values if absorb_singular_values in ("left", "right") else "both"
    U2, V2 = _absorb(U, S, V, how)
    return U2, V2, truncated_evals


def _qr_device(d, nrow):
    rows, cols, m, n = _matrix_modes(d, nrow)
    k = min(m, n)
    q = torch.empty((m, k), dtype=d.dtype, device=d.device)
    r = torch.empty((k, n), dtype=d.dtype, device=d.device)
    if k == 0:
        return q, r
    lib = L.lib()
    ws = L.workspace(lib.tn_qr_workspace(m, n, L.dtype_code(d)))
    L.check(lib.tn_qr(L.ptr(d), ctypes.byref(rows), ctypes.byref(cols), m, n, L.ptr(q), L.ptr(r),
                      L.dtype_code(d), L.ptr(ws), ws.numel(), L.stream_ptr()))
    return q, r


def tensor_qr(tensor, row_labels, qr_label="qr_"):
    """tensor.py:962-1055 -- blocked Householder QR on the GPU."""
    if not isinstance(row_labels, list):
        row_labels = [row_labels]
    t, nrow = _row_view(tensor, row_labels, move_all=True)
    col_labels = [x for x in t.labels if x not in row_labels]
    old_shape = t.shape
    q, r = _qr_device(t.data, nrow)
    k = q.shape[1]
    Q = Tensor._wrap(q.view(tuple(old_shape[:len(row_labels)]) + (k,)), row_labels + [qr_label + "in"])
    R = Tensor._wrap(r.view((k,) + tuple(old_shape[len(row_labels):])), [qr_label + "out"] + col_labels)
    return Q, R


def tensor_lq(tensor, row_labels, lq_label="lq_"):
    """tensor.py:1058-1109 (QR of the column labels, relabelled)."""
    col_labels = [x for x in tensor.labels if x not in row_labels]
    tmp = lbl.unique_label()
    Q, Lm = tensor_qr(tensor, col_labels, qr_label=tmp)
    Q.replace_label(tmp + "in", lq_label + "out")
    Lm.replace_label(tmp + "out", lq_label + "in")
    return Lm, Q


def _cholqr_device(d, nrow):
    """Two-pass Cholesky QR (tn_cholqr) of the (rows | cols) view of the FP64
    tensor d with m >= n: (q m x n, r n x n) on the device, or None when the
    view is wide, not FP64 or too ill-conditioned for Cholesky QR."""
    rows, cols, m, n = _matrix_modes(d, nrow)
    if d.dtype != torch.float64 or m < n or n == 0:
        return None
    q = torch.empty((m, n), dtype=d.dtype, device=d.device)
    r = torch.empty((n, n), dtype=d.dtype, device=d.device)
    lib = L.lib()
    ws = L.workspace(lib.tn_cholqr_workspace(m, n))
    orth = ctypes.c_double(0.0)
    rc = lib.tn_cholqr(L.ptr(d), ctypes.byref(rows), ctypes.byref(cols), m, n, L.ptr(q), L.ptr(r),
                       0.0, ctypes.byref(orth), L.ptr(ws), ws.numel(), L.stream_ptr())
    if rc == L.TN_ERR_NOCONV:
        return None
    L.check(rc)
    return q, r


def _lq_rowspace(tensor, row_labels, lq_label="lq_"):
    """tensor_lq's (L, Q) for a caller that uses only the row space of Q and
    L Q = tensor: variational_compress (onedim_core.py:560-631), where Q
    becomes the site and L is absorbed into the neighbour.  Factors by
    two-pass Cholesky QR on the GPU, so each row of Q (column of L) may differ
    in sign from Householder's; blocks too ill-conditioned for it, wide ones
    and non-FP64 ones go through tensor_lq.  Labels and shapes are
    tensor_lq's (tensor.py:1058-1109)."""
    col_labels = [x for x in tensor.labels if x not in row_labels]
    t, nrow = _row_view(tensor, col_labels, move_all=True)
    res = _cholqr_device(t.data, nrow)
    if res is None:
        return tensor_lq(tensor, row_labels, lq_label)
    q, r = res
    k = q.shape[1]
    shape = t.shape
    Q = Tensor._wrap(q.view(tuple(shape[:nrow]) + (k,)), list(t.labels[:nrow]) + [lq_label + "out"])
    Lm = Tensor._wrap(r.view((k,) + tuple(shape[nrow:])), [lq_label + "in"] + list(t.labels[nrow:]))
    return Lm, Q


def conjugate(tensor):
    t = tensor.copy()
    t.conjugate()
    return t
