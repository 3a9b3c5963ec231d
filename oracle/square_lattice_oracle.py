"""CPU FP64 oracle for tncontract/twodim/square_lattice.py.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product package ``tncontract_amd`` never imports it.

NumPy restatement of the square-lattice tensor networks (PEPS, PEPO and
their inner/outer products) on top of the one-dimensional oracle.  How each
function is pinned (checks in ``tests/test_oracle_square_lattice.py``):
  * exact_contract: ``np.einsum`` over the whole lattice with explicit bond
    indices (brute force, lattices up to 3 x 4), also for ``until_column``;
  * mps_contract (svd and variational): equal to exact_contract when chi is at
    least the exact boundary bond dimension; ``until_column`` and
    ``return_all_columns`` against the partial exact contractions;
  * inner_product_peps / outer_product_peps / SquareLatticePEPS.outer_product
    / SquareLatticePEPO.trace / apply_pepo_to_peps: product-state closed forms
    (bond dimension 1: the product of single-site overlaps), trace of the
    outer product equal to the inner product, identity PEPO leaves the PEPS;
  * fliplr: the mirrored lattice contracts to the same scalar;
  * can_contract: mismatched bonds are reported.
"""
from __future__ import annotations

import numpy as np

from oracle import tn_oracle as o


def _object_grid(rows):
    grid = np.empty((len(rows), len(rows[0]) if rows else 0), dtype=object)
    for i, row in enumerate(rows):
        for j, x in enumerate(row):
            grid[i, j] = x
    return grid


class OSquareTN:
    """square_lattice.py:14-228 -- 2D object array of tensors with up, right,
    down and left indices (dimension-1 dummies added where absent)."""

    def __init__(self, tensors, up_label="up", right_label="right", down_label="down",
                 left_label="left", copy_data=True):
        self.up_label, self.right_label = up_label, right_label
        self.down_label, self.left_label = down_label, left_label
        rows = [[x.copy() if copy_data else x for x in row] for row in tensors]
        self.data = _object_grid(rows)
        for _, x in np.ndenumerate(self.data):
            for lab in (left_label, right_label, up_label, down_label):
                if lab not in x.labels:
                    x.add_dummy_index(lab)

    def __iter__(self):
        return iter(self.data)

    def __len__(self):
        return len(self.data)

    def __getitem__(self, key):
        return self.data[key]

    def __setitem__(self, key, value):
        self.data[key] = value

    def copy(self):
        return OSquareTN(self.data, self.up_label, self.right_label, self.down_label,
                         self.left_label, copy_data=True)

    @property
    def shape(self):
        return self.data.shape

    def is_left_right_periodic(self):
        """square_lattice.py:82-89"""
        return any(x.index_dimension(self.left_label) > 1 for x in self[:, 0])

    def can_contract(self):
        """square_lattice.py:91-113 -- interior bonds (i, j >= 1) only."""
        rows, cols = self.shape
        bad_v, bad_h = [], []
        for i in range(1, rows):
            for j in range(1, cols):
                if self[i, j].index_dimension(self.up_label) != \
                        self[i - 1, j].index_dimension(self.down_label):
                    bad_v.append((i, j))
                if self[i, j].index_dimension(self.left_label) != \
                        self[i, j - 1].index_dimension(self.right_label):
                    bad_h.append((i, j))
        return not bad_v and not bad_h

    def exact_contract(self, until_column=-1):
        """square_lattice.py:115-134 -- column 0 as an MPS contracted along its
        bonds, then each further column (an MPO) absorbed through its left
        indices; returns early after `until_column`."""
        rows, cols = self.shape
        C = o.contract_virtual_indices(column_to_mpo(self, 0))
        for i in range(1, cols):
            if i == until_column + 1:
                return C
            mpo = column_to_mpo(self, i)
            C = o.contract_multi_index_tensor_with_one_dim_array(C, mpo, self.right_label,
                                                                 self.left_label)
            C.remove_all_dummy_indices([self.left_label, self.up_label, self.down_label])
        return C

    def mps_contract(self, chi, compression_type="svd", until_column=-1, max_iter=10,
                     tolerance=1e-14, return_all_columns=False):
        """square_lattice.py:136-203 -- boundary-MPS contraction left to right:
        MPO application + svd (right-canonical, normalised at site 0) or
        variational (left-canonical, normalised at the last site) compression
        per column, the removed norms accumulated in long double; the last
        column is contracted exactly against the boundary MPS."""
        nrows, ncols = self.shape
        columns = [] if return_all_columns else None
        norm = np.longdouble(1)
        compressed = None
        for col in range(ncols - 1):
            if col == 0:
                to_compress = column_to_mpo(self, 0)
            else:
                to_compress = o.contract_mps_mpo(compressed, column_to_mpo(self, col))
            if compression_type == "svd":
                compressed = o.svd_compress_mps(to_compress, chi, normalise=False,
                                                threshold=tolerance)
                mps_norm = compressed.norm(canonical_form="right")
                if mps_norm == 0.0:
                    return 0.0
                compressed[0].data = compressed[0].data / mps_norm
                norm *= mps_norm
            elif compression_type == "variational":
                compressed = to_compress.variational_compress(chi, max_iter=max_iter,
                                                              tolerance=tolerance)
                mps_norm = compressed.norm(canonical_form="left")
                if mps_norm == 0.0:
                    return 0.0
                compressed[-1].data = compressed[-1].data / mps_norm
                norm *= mps_norm
            if return_all_columns:
                c = compressed.copy()
                c[0].data *= norm
                columns.append(c)
            if col == until_column:
                if return_all_columns:
                    return columns
                if compression_type == "svd":
                    compressed[0].data = np.longdouble(compressed[0].data)
                    compressed[0].data *= norm
                    return compressed
                if compression_type == "variational":
                    compressed[-1].data *= norm
                    return compressed
        final = column_to_mpo(self, ncols - 1)
        full = o.inner_product_mps(compressed, final, return_whole_tensor=True,
                                   complex_conjugate_bra=False) * norm
        if return_all_columns:
            columns.append(full)
            return columns
        return full

    def col_to_1D_array(self, col):
        """square_lattice.py:205-215"""
        return o._O1D(list(self[:, col]), self.up_label, self.down_label)

    def fliplr(self):
        """square_lattice.py:217-228 -- mirrored columns; left/right label
        attributes swapped, tensor labels untouched."""
        m = self.copy()
        m.data = np.fliplr(m.data)
        m.right_label, m.left_label = self.left_label, self.right_label
        return m


class OPEPS(OSquareTN):
    """square_lattice.py:233-279"""

    def __init__(self, tensors, up_label="up", right_label="right", down_label="down",
                 left_label="left", phys_label="phys", copy_data=True):
        super().__init__(tensors, up_label, right_label, down_label, left_label, copy_data)
        self.phys_label = phys_label

    def copy(self):
        return OPEPS(self.data, self.up_label, self.right_label, self.down_label,
                     self.left_label, self.phys_label, copy_data=True)

    def outer_product(self, physin_label="physin", physout_label="physout"):
        """square_lattice.py:250-276 -- site-wise outer product with itself,
        first copy's phys -> physin, second's -> physout, virtual pairs fused."""
        grid = []
        for r in range(self.shape[0]):
            row = []
            for c in range(self.shape[1]):
                t = o.contract(self[r, c], self[r, c], [], [])
                t.labels[t.labels.index(self.phys_label)] = physin_label
                t.labels[t.labels.index(self.phys_label)] = physout_label
                t.consolidate_indices(labels=[self.left_label, self.right_label, self.up_label,
                                              self.down_label])
                row.append(t)
            grid.append(row)
        return OPEPO(grid, up_label=self.up_label, down_label=self.down_label,
                     right_label=self.right_label, left_label=self.left_label,
                     physin_label=physin_label, physout_label=physout_label)

    density_operator = outer_product


class OPEPO(OSquareTN):
    """square_lattice.py:344-375"""

    def __init__(self, tensors, up_label="up", right_label="right", down_label="down",
                 left_label="left", physin_label="physin", physout_label="physout",
                 copy_data=True):
        super().__init__(tensors, up_label, right_label, down_label, left_label, copy_data)
        self.physin_label = physin_label
        self.physout_label = physout_label

    def copy(self):
        return OPEPO(self.data, self.up_label, self.right_label, self.down_label,
                     self.left_label, self.physin_label, self.physout_label, copy_data=True)

    def trace(self):
        """square_lattice.py:363-375"""
        grid = []
        for i in range(self.shape[0]):
            row = []
            for j in range(self.shape[1]):
                t = self[i, j].copy()
                t.trace(self.physin_label, self.physout_label)
                row.append(t)
            grid.append(row)
        return OSquareTN(grid, up_label=self.up_label, down_label=self.down_label,
                         right_label=self.right_label, left_label=self.left_label)


def inner_product_peps(peps_ket, peps_bra, exact_contract="True", complex_conjugate_bra=True,
                       compression_type="svd", chi=2, max_iter=10, tolerance=1e-14,
                       contract_virtual=True):
    """square_lattice.py:281-304 -- contract each bra/ket site pair over
    "phys", fuse the virtual pairs, contract the resulting lattice."""
    grid = []
    for i in range(peps_ket.shape[0]):
        row = []
        for j in range(peps_ket.shape[1]):
            t = o.conjugate(peps_bra[i, j])["phys"] * peps_ket[i, j]["phys"]
            t.consolidate_indices()
            row.append(t)
        grid.append(row)
    ip = OSquareTN(grid)
    if not contract_virtual:
        return ip
    if exact_contract:
        return ip.exact_contract()
    return ip.mps_contract(chi, compression_type=compression_type, max_iter=max_iter,
                           tolerance=tolerance)


def outer_product_peps(peps1, peps2, physin_label="physin", physout_label="physout"):
    """square_lattice.py:306-342 -- |peps1><peps2| as a PEPO."""
    if peps1.shape != peps2.shape:
        raise ValueError("Peps input do not have same dimension.")
    grid = []
    for r in range(peps1.shape[0]):
        row = []
        for c in range(peps1.shape[1]):
            t = o.contract(peps1[r, c], o.conjugate(peps2[r, c]), [], [])
            t.labels[t.labels.index(peps1.phys_label)] = physout_label
            t.labels[t.labels.index(peps2.phys_label)] = physin_label
            t.consolidate_indices(labels=[peps1.left_label, peps1.right_label, peps1.up_label,
                                          peps1.down_label])
            row.append(t)
        grid.append(row)
    return OPEPO(grid, up_label=peps1.up_label, down_label=peps1.down_label,
                 right_label=peps1.right_label, left_label=peps1.left_label,
                 physin_label=physin_label, physout_label=physout_label)


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
    return OPEPS(grid, up_label=peps.up_label, right_label=peps.right_label,
                 down_label=peps.down_label, left_label=peps.left_label,
                 phys_label=peps.phys_label)


def column_to_mpo(square_tn, col):
    """square_lattice.py:395-417 -- first/last column: an MPS along up/down
    whose physical index is the right/left bond (the outer dummy removed);
    interior columns: an MPO with physout = right and physin = left."""
    col_data = list(square_tn[:, col])
    ncols = square_tn.shape[1]
    if col == 0 or col == ncols - 1:
        if col == 0:
            mps = o.OMPS(col_data, square_tn.up_label, square_tn.down_label,
                         square_tn.right_label)
            for x in mps.data:
                x.remove_all_dummy_indices(square_tn.left_label)
        else:
            mps = o.OMPS(col_data, square_tn.up_label, square_tn.down_label,
                         square_tn.left_label)
            for x in mps.data:
                x.remove_all_dummy_indices(square_tn.right_label)
        return mps
    return o.OMPO(col_data, square_tn.up_label, square_tn.down_label, square_tn.right_label,
                  square_tn.left_label)
