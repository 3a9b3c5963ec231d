"""Square-lattice PEPS / PEPO networks on the GPU (tncontract/twodim/
square_lattice.py).

Same classes, functions, labels and contraction order as the reference: the
lattice is a 2D object array of device ``Tensor``s with up/right/down/left
indices; columns become MPS / MPOs and every contraction, compression and
overlap runs through libtncontract_amd (MFMA contractions, Householder QR and
Jacobi SVD).  The boundary-MPS norms removed during ``mps_contract`` are kept
on the host in ``np.longdouble`` exactly as the reference keeps them; device
data is FP64, so a result whose value lies outside the FP64 range raises
OverflowError where the reference would return a long-double array.
"""
from __future__ import annotations

import numpy as np

from tncontract_amd import tensor as tsr
from tncontract_amd import onedim as od

__all__ = ["SquareLatticeTensorNetwork", "SquareLatticePEPS", "SquareLatticePEPO",
           "inner_product_peps", "outer_product_peps", "apply_pepo_to_peps", "column_to_mpo"]


def _object_grid(rows):
    grid = np.empty((len(rows), len(rows[0]) if rows else 0), dtype=object)
    for i, row in enumerate(rows):
        for j, x in enumerate(row):
            grid[i, j] = x
    return grid


def _as_rows(tensors):
    if isinstance(tensors, np.ndarray):
        return [list(tensors[i, :]) for i in range(tensors.shape[0])]
    return [list(r) for r in tensors]


def _fp64_factor(norm):
    """A long-double norm as an FP64 scale factor (OverflowError when the
    FP64 range cannot hold it)."""
    f = float(norm)
    if not np.isfinite(f) or (f == 0.0 and norm != 0):
        raise OverflowError("boundary-MPS norm %s lies outside the FP64 range of device data"
                            % repr(norm))
    return f


def _times_longdouble(t, norm):
    """Tensor t times a long-double factor, multiplied on the host in long
    double (t is the small final contraction) and returned as FP64."""
    v = t.numpy().astype(np.longdouble) * norm
    with np.errstate(over="ignore", under="ignore"):
        f = v.astype(np.float64)
    if not np.all(np.isfinite(f)) or np.any((f == 0.0) & (v != 0)):
        raise OverflowError("contraction value lies outside the FP64 range of device data")
    return tsr.Tensor(f, list(t.labels))


class SquareLatticeTensorNetwork:
    """square_lattice.py:14-228"""

    def __init__(self, tensors, up_label="up", right_label="right", down_label="down",
                 left_label="left", copy_data=True):
        self.up_label = up_label
        self.right_label = right_label
        self.down_label = down_label
        self.left_label = left_label
        rows = _as_rows(tensors)
        if copy_data:
            rows = [[x.copy() for x in row] for row in rows]
        self.data = _object_grid(rows)
        for _, x in np.ndenumerate(self.data):
            if left_label not in x.labels:
                x.add_dummy_index(left_label)
            if right_label not in x.labels:
                x.add_dummy_index(right_label)
            if up_label not in x.labels:
                x.add_dummy_index(up_label)
            if down_label not in x.labels:
                x.add_dummy_index(down_label)

    def __iter__(self):
        return self.data.__iter__()

    def __len__(self):
        return self.data.__len__()

    def __getitem__(self, key):
        return self.data.__getitem__(key)

    def __setitem__(self, key, value):
        self.data.__setitem__(key, value)

    def copy(self):
        return SquareLatticeTensorNetwork(self.data, up_label=self.up_label,
                                          right_label=self.right_label,
                                          down_label=self.down_label, left_label=self.left_label,
                                          copy_data=True)

    @property
    def shape(self):
        return self.data.shape

    def is_left_right_periodic(self):
        """square_lattice.py:82-89"""
        for x in self[:, 0]:
            if x.index_dimension(self.left_label) > 1:
                return True
        return False

    def can_contract(self):
        """square_lattice.py:91-113 (interior bonds, i, j >= 1)."""
        rows, cols = self.data.shape
        vert, horiz = [], []
        for i in range(1, rows):
            for j in range(1, cols):
                if self[i, j].index_dimension(self.up_label) != \
                        self[i - 1, j].index_dimension(self.down_label):
                    vert.append((i, j))
                if self[i, j].index_dimension(self.left_label) != \
                        self[i, j - 1].index_dimension(self.right_label):
                    horiz.append((i, j))
        if not vert and not horiz:
            return True
        print("Unmatched bonds found between the following sites:")
        for k in vert:
            print("(" + str(k[0] - 1) + ", " + str(k[1]) + ")" + " and " + str(k))
        for k in horiz:
            print("(" + str(k[0]) + ", " + str(k[1] - 1) + ")" + " and " + str(k))
        return False

    def exact_contract(self, until_column=-1):
        """square_lattice.py:115-134: column 0 contracted along its bonds, then
        each further column (an MPO) absorbed through its left indices."""
        rows, cols = self.data.shape
        C = od.contract_virtual_indices(column_to_mpo(self, 0))
        for i in range(1, cols):
            if i == until_column + 1:
                return C
            mpo = column_to_mpo(self, i)
            C = od.contract_multi_index_tensor_with_one_dim_array(C, mpo, self.right_label,
                                                                  self.left_label)
            C.remove_all_dummy_indices([self.left_label, self.up_label, self.down_label])
        return C

    def mps_contract(self, chi, compression_type="svd", until_column=-1, max_iter=10,
                     tolerance=1e-14, return_all_columns=False):
        """square_lattice.py:136-203: boundary MPS swept left to right; each
        column's MPO is applied (contract_mps_mpo) and the result compressed
        (svd_compress_mps: right canonical, normalised at site 0; or
        variational_compress: left canonical, normalised at the last site),
        the removed norms accumulated in long double on the host; the last
        column is contracted exactly against the boundary MPS."""
        nrows, ncols = self.shape
        columns = [] if return_all_columns else None
        norm = np.longdouble(1)
        compressed = None
        for col in range(ncols - 1):
            if col == 0:
                to_compress = column_to_mpo(self, 0)
            else:
                to_compress = od.contract_mps_mpo(compressed, column_to_mpo(self, col))
            if compression_type == "svd":
                compressed = od.svd_compress_mps(to_compress, chi, normalise=False,
                                                 threshold=tolerance)
                mps_norm = compressed.norm(canonical_form="right")
                if mps_norm == 0.0:
                    return 0.0
                compressed[0].scale_(1.0 / mps_norm)
                norm *= mps_norm
            elif compression_type == "variational":
                compressed = to_compress.variational_compress(chi, max_iter=max_iter,
                                                              tolerance=tolerance)
                mps_norm = compressed.norm(canonical_form="left")
                if mps_norm == 0.0:
                    return 0.0
                compressed[-1].scale_(1.0 / mps_norm)
                norm *= mps_norm
            if return_all_columns:
                c = compressed.copy()
                c[0].scale_(_fp64_factor(norm))
                columns.append(c)
            if col == until_column:
                if return_all_columns:
                    return columns
                if compression_type == "svd":
                    compressed[0].scale_(_fp64_factor(norm))
                    return compressed
                if compression_type == "variational":
                    compressed[-1].scale_(_fp64_factor(norm))
                    return compressed
        final = column_to_mpo(self, ncols - 1)
        full = _times_longdouble(od.inner_product_mps(compressed, final, return_whole_tensor=True,
                                                      complex_conjugate_bra=False), norm)
        if return_all_columns:
            columns.append(full)
            return columns
        return full

    def col_to_1D_array(self, col):
        """square_lattice.py:205-215"""
        return od.OneDimensionalTensorNetwork(list(self[:, col]), left_label=self.up_label,
                                              right_label=self.down_label)

    def fliplr(self):
        """square_lattice.py:217-228: mirrored columns, left/right label
        attributes swapped (tensor labels untouched)."""
        mirror = self.copy()
        mirror.data = np.fliplr(mirror.data)
        mirror.right_label = self.left_label
        mirror.left_label = self.right_label
        return mirror


class SquareLatticePEPS(SquareLatticeTensorNetwork):
    """square_lattice.py:233-279"""

    def __init__(self, tensors, up_label="up", right_label="right", down_label="down",
                 left_label="left", phys_label="phys", copy_data=True):
        SquareLatticeTensorNetwork.__init__(self, tensors, up_label, right_label, down_label,
                                            left_label, copy_data=copy_data)
        self.phys_label = phys_label

    def copy(self):
        return SquareLatticePEPS(self.data, up_label=self.up_label, right_label=self.right_label,
                                 down_label=self.down_label, left_label=self.left_label,
                                 phys_label=self.phys_label, copy_data=True)

    def outer_product(self, physin_label="physin", physout_label="physout"):
        """square_lattice.py:250-276: site-wise outer product with itself (the
        first copy's phys becomes physin, the second's physout), virtual pairs
        fused."""
        grid = []
        for row in range(self.shape[0]):
            new_row = []
            for col in range(self.shape[1]):
                outer = tsr.contract(self[row, col], self[row, col], [], [])
                outer.labels[outer.labels.index(self.phys_label)] = physin_label
                outer.labels[outer.labels.index(self.phys_label)] = physout_label
                outer.consolidate_indices(labels=[self.left_label, self.right_label,
                                                  self.up_label, self.down_label])
                new_row.append(outer)
            grid.append(new_row)
        return SquareLatticePEPO(grid, up_label=self.up_label, down_label=self.down_label,
                                 right_label=self.right_label, left_label=self.left_label,
                                 physin_label=physin_label, physout_label=physout_label)

    density_operator = outer_product


def inner_product_peps(peps_ket, peps_bra, exact_contract="True", complex_conjugate_bra=True,
                       compression_type="svd", chi=2, max_iter=10, tolerance=1e-14,
                       contract_virtual=True):
    """square_lattice.py:281-304: bra/ket site pairs contracted over "phys",
    virtual pairs fused, the resulting lattice contracted exactly or by
    boundary MPS."""
    grid = []
    for i in range(peps_ket.shape[0]):
        row = []
        for j in range(peps_ket.shape[1]):
            t = tsr.conjugate(peps_bra[i, j])["phys"] * peps_ket[i, j]["phys"]
            t.consolidate_indices()
            row.append(t)
        grid.append(row)
    ip = SquareLatticeTensorNetwork(grid, copy_data=False)
    if not contract_virtual:
        return ip
    if exact_contract:
        return ip.exact_contract()
    return ip.mps_contract(chi, compression_type=compression_type, max_iter=max_iter,
                           tolerance=tolerance)


def outer_product_peps(peps1, peps2, physin_label="physin", physout_label="physout"):
    """square_lattice.py:306-342: |peps1><peps2| as a PEPO (physout from
    peps1, physin from peps2)."""
    if peps1.shape != peps2.shape:
        raise ValueError("Peps input do not have same dimension.")
    grid = []
    for row in range(peps1.shape[0]):
        new_row = []
        for col in range(peps1.shape[1]):
            outer = tsr.contract(peps1[row, col], tsr.conjugate(peps2[row, col]), [], [])
            outer.labels[outer.labels.index(peps1.phys_label)] = physout_label
            outer.labels[outer.labels.index(peps2.phys_label)] = physin_label
            outer.consolidate_indices(labels=[peps1.left_label, peps1.right_label,
                                              peps1.up_label, peps1.down_label])
            new_row.append(outer)
        grid.append(new_row)
    return SquareLatticePEPO(grid, up_label=peps1.up_label, down_label=peps1.down_label,
                             right_label=peps1.right_label, left_label=peps1.left_label,
                             physin_label=physin_label, physout_label=physout_label)


class SquareLatticePEPO(SquareLatticeTensorNetwork):
    """square_lattice.py:344-375"""

    def __init__(self, tensors, up_label="up", right_label="right", down_label="down",
                 left_label="left", physin_label="physin", physout_label="physout",
                 copy_data=True):
        SquareLatticeTensorNetwork.__init__(self, tensors, up_label, right_label, down_label,
                                            left_label, copy_data=copy_data)
        self.physin_label = physin_label
        self.physout_label = physout_label

    def copy(self):
        return SquareLatticePEPO(self.data, up_label=self.up_label, right_label=self.right_label,
                                 down_label=self.down_label, left_label=self.left_label,
                                 physin_label=self.physin_label,
                                 physout_label=self.physout_label, copy_data=True)

    def trace(self):
        """square_lattice.py:363-375: physin traced with physout on every
        site (one device trace each)."""
        grid = []
        for i in range(self.shape[0]):
            row = []
            for j in range(self.shape[1]):
                t = self[i, j].copy()
                t.trace(self.physin_label, self.physout_label)
                row.append(t)
            grid.append(row)
        return SquareLatticeTensorNetwork(grid, up_label=self.up_label,
                                          down_label=self.down_label,
                                          right_label=self.right_label,
                                          left_label=self.left_label, copy_data=False)


def apply_pepo_to_peps(peps, pepo):
    """square_lattice.py:377-393"""
    grid = []
    for i in range(peps.shape[0]):
        row = []
        for j in range(peps.shape[1]):
            t = peps[i, j][peps.phys_label] * pepo[i, j][pepo.physin_label]
            t.replace_label(pepo.physout_label, peps.phys_label)
            t.consolidate_indices()
            row.append(t)
        grid.append(row)
    return SquareLatticePEPS(grid, up_label=peps.up_label, right_label=peps.right_label,
                             down_label=peps.down_label, left_label=peps.left_label,
                             phys_label=peps.phys_label, copy_data=False)


def column_to_mpo(square_tn, col):
    """square_lattice.py:395-417: the first (last) column as an MPS along
    up/down whose physical index is the right (left) bond, the outer dummy
    index removed; interior columns as an MPO (physout = right, physin =
    left)."""
    col_data = list(square_tn[:, col])
    if col == 0 or col == square_tn.shape[1] - 1:
        if col == 0:
            mps = od.MatrixProductState(col_data, square_tn.up_label, square_tn.down_label,
                                        square_tn.right_label)
            for x in mps.data:
                x.remove_all_dummy_indices(square_tn.left_label)
        else:
            mps = od.MatrixProductState(col_data, square_tn.up_label, square_tn.down_label,
                                        square_tn.left_label)
            for x in mps.data:
                x.remove_all_dummy_indices(square_tn.right_label)
        return mps
    return od.MatrixProductOperator(col_data, square_tn.up_label, square_tn.down_label,
                                    square_tn.right_label, square_tn.left_label)
