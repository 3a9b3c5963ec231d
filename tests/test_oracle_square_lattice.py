"""Pins of the square-lattice oracle (tncontract/twodim/square_lattice.py) by
brute force and closed forms: np.einsum over the whole lattice, product-state
overlaps, trace of |a><b| = <b|a>, boundary-MPS contraction exact at large chi."""
import numpy as np
import pytest

from oracle import tn_oracle as o
from oracle import square_lattice_oracle as sq

VIRT = ["up", "right", "down", "left"]


def random_grid(rows, cols, D, rng, phys=None, drop_boundary=False):
    """Tensors labelled (phys?, up, right, down, left); boundary bonds have
    dimension 1 (or are left out when drop_boundary, so the container adds
    them)."""
    grid = []
    for i in range(rows):
        row = []
        for j in range(cols):
            dims = {"up": 1 if i == 0 else D, "down": 1 if i == rows - 1 else D,
                    "left": 1 if j == 0 else D, "right": 1 if j == cols - 1 else D}
            labels = [x for x in VIRT if not (drop_boundary and dims[x] == 1)]
            shape = [dims[x] for x in labels]
            if phys:
                labels, shape = ["phys"] + labels, [phys] + shape
            row.append(o.OTensor(rng.standard_normal(shape), labels))
        grid.append(row)
    return grid


def bond_ids(rows, cols):
    """einsum index of every bond: (i, j, label) -> int.  All boundary bonds
    (dimension 1) share index 0; interior bonds are numbered from 1."""
    ids, nxt = {}, [1]

    def new():
        nxt[0] += 1
        return nxt[0] - 1
    for i in range(rows):
        for j in range(cols):
            ids[(i, j, "right")] = 0 if j == cols - 1 else ids.setdefault((i, j + 1, "left"), new())
            ids.setdefault((i, j, "left"), 0)
            ids[(i, j, "down")] = 0 if i == rows - 1 else ids.setdefault((i + 1, j, "up"), new())
            ids.setdefault((i, j, "up"), 0)
    return ids, nxt[0]


def brute_force(tn, open_right_col=None):
    """Full contraction of a square network (four virtual labels per tensor)
    by np.einsum; with open_right_col=k only columns 0..k, leaving their
    right bonds of column k open (in row order)."""
    rows, cols = tn.shape
    cmax = cols if open_right_col is None else open_right_col + 1
    ids, _ = bond_ids(rows, cols)
    ops = []
    for i in range(rows):
        for j in range(cmax):
            t = tn[i, j]
            ops += [t.data, [ids[(i, j, lab)] for lab in t.labels]]
    out = [] if open_right_col is None else [ids[(i, cmax - 1, "right")] for i in range(rows)]
    return np.einsum(*ops, out, optimize=True)


def square(grid):
    return sq.OSquareTN(grid)


@pytest.mark.parametrize("rows,cols,D", [(2, 3, 3), (3, 3, 2), (3, 4, 2), (1, 4, 3), (4, 2, 2)])
def test_exact_contract_brute_force(rows, cols, D, rng):
    tn = square(random_grid(rows, cols, D, rng))
    full = tn.exact_contract()
    np.testing.assert_allclose(np.squeeze(full.data), brute_force(tn), rtol=1e-12)
    for k in range(cols - 1):
        part = tn.exact_contract(until_column=k)
        assert all(x == "right" for x in part.labels if part.index_dimension(x) > 1)
        got = part.data.reshape([part.data.shape[a] for a in range(part.data.ndim)])
        ref = brute_force(tn, open_right_col=k)
        np.testing.assert_allclose(got.reshape(ref.shape), ref, rtol=1e-12, atol=1e-12)


def test_exact_contract_adds_missing_boundary_indices(rng):
    grid = random_grid(3, 3, 2, rng, drop_boundary=True)
    tn = square(grid)
    assert all(set(x.labels) == set(VIRT) for _, x in np.ndenumerate(tn.data))
    np.testing.assert_allclose(np.squeeze(tn.exact_contract().data), brute_force(tn), rtol=1e-12)


@pytest.mark.parametrize("ctype", ["svd", "variational"])
def test_mps_contract_exact_at_large_chi(ctype, rng):
    rows, cols, D = 4, 4, 2
    tn = square(random_grid(rows, cols, D, rng))
    ref = brute_force(tn)
    got = tn.mps_contract(D ** rows, compression_type=ctype, tolerance=1e-10, max_iter=20)
    np.testing.assert_allclose(float(np.squeeze(got.data)), ref, rtol=1e-10)
    cols_out = tn.mps_contract(D ** rows, compression_type=ctype, tolerance=1e-10, max_iter=20,
                               return_all_columns=True)
    assert len(cols_out) == cols
    np.testing.assert_allclose(float(np.squeeze(cols_out[-1].data)), ref, rtol=1e-10)
    from oracle import models as om
    for k, m in enumerate(cols_out[:-1]):
        part = brute_force(tn, open_right_col=k)
        np.testing.assert_allclose(om.mps_to_dense(m).reshape(part.shape), part,
                                   rtol=1e-9, atol=1e-10 * np.abs(part).max())
    # until_column: the boundary MPS after column k, its norm restored
    m = tn.mps_contract(D ** rows, compression_type=ctype, tolerance=1e-10, max_iter=20,
                        until_column=1)
    part = brute_force(tn, open_right_col=1)
    dense = om.mps_to_dense(o.OMPS([o.OTensor(np.asarray(t.data, dtype=np.float64), t.labels)
                                    for t in m], m.left_label, m.right_label, m.phys_label))
    np.testing.assert_allclose(dense.reshape(part.shape), part, rtol=1e-9,
                               atol=1e-10 * np.abs(part).max())
    if ctype == "svd":
        assert m[0].data.dtype == np.longdouble


def test_mps_contract_truncated_is_close_and_zero_norm(rng):
    tn = square(random_grid(4, 5, 3, rng))
    ref = brute_force(tn)
    approx = float(np.squeeze(tn.mps_contract(6).data))
    assert np.isfinite(approx)
    exact = float(np.squeeze(tn.mps_contract(81).data))
    np.testing.assert_allclose(exact, ref, rtol=1e-9)
    z = square(random_grid(3, 3, 2, rng))
    for x in z[:, 0]:
        x.data[:] = 0
    assert z.mps_contract(4) == 0.0


def product_peps(rows, cols, d, rng):
    vecs = rng.standard_normal((rows, cols, d))
    grid = [[o.OTensor(vecs[i, j].reshape(d, 1, 1, 1, 1), ["phys"] + VIRT) for j in range(cols)]
            for i in range(rows)]
    return sq.OPEPS(grid), vecs


def test_inner_product_product_state_closed_form(rng):
    a, va = product_peps(3, 4, 3, rng)
    b, vb = product_peps(3, 4, 3, rng)
    closed = np.prod(np.sum(va * vb, axis=2))
    for exact in (True, False):
        ip = sq.inner_product_peps(a, b, exact_contract=exact, chi=2)
        np.testing.assert_allclose(float(np.squeeze(ip.data)), closed, rtol=1e-12)
    net = sq.inner_product_peps(a, b, contract_virtual=False)
    assert isinstance(net, sq.OSquareTN) and net.shape == (3, 4)


def test_inner_product_double_layer_brute_force(rng):
    rows, cols, D, d = 3, 3, 2, 2
    a = sq.OPEPS(random_grid(rows, cols, D, rng, phys=d))
    b = sq.OPEPS(random_grid(rows, cols, D, rng, phys=d))
    ip = float(np.squeeze(sq.inner_product_peps(a, b).data))
    # brute force on the two layers with separate bonds, summing phys per site
    ids_a, n = bond_ids(rows, cols)
    ids_b = {k: (v + n if v else 0) for k, v in ids_a.items()}
    ops, p = [], 2 * n
    for i in range(rows):
        for j in range(cols):
            for t, ids in ((a[i, j], ids_a), (b[i, j], ids_b)):
                ops += [t.data, [p if lab == "phys" else ids[(i, j, lab)] for lab in t.labels]]
            p += 1
    np.testing.assert_allclose(ip, np.einsum(*ops, [], optimize=True), rtol=1e-12)
    # boundary-MPS route at exact chi, both compressions
    for ctype in ("svd", "variational"):
        got = sq.inner_product_peps(a, b, exact_contract=False, chi=(D * D) ** rows,
                                    compression_type=ctype, tolerance=1e-10, max_iter=20)
        np.testing.assert_allclose(float(np.squeeze(got.data)), ip, rtol=1e-10)


def test_outer_product_trace_equals_inner_product(rng):
    a = sq.OPEPS(random_grid(3, 3, 2, rng, phys=2))
    b = sq.OPEPS(random_grid(3, 3, 2, rng, phys=2))
    ab = sq.outer_product_peps(a, b)
    assert isinstance(ab, sq.OPEPO)
    assert set(ab[1, 1].labels) == set(VIRT + ["physin", "physout"])
    assert ab[1, 1].index_dimension("up") == 4
    tr = float(np.squeeze(ab.trace().exact_contract().data))
    np.testing.assert_allclose(tr, float(np.squeeze(sq.inner_product_peps(a, b).data)), rtol=1e-12)
    rho = a.outer_product()
    assert a.density_operator.__func__ is a.outer_product.__func__
    np.testing.assert_allclose(float(np.squeeze(rho.trace().exact_contract().data)),
                               float(np.squeeze(sq.inner_product_peps(a, a).data)), rtol=1e-12)
    with pytest.raises(ValueError):
        sq.outer_product_peps(a, sq.OPEPS(random_grid(2, 3, 2, rng, phys=2)))


def test_apply_identity_pepo_and_pepo_expectation(rng):
    a = sq.OPEPS(random_grid(2, 3, 2, rng, phys=2))
    eye = [[o.OTensor(np.eye(2).reshape(2, 2, 1, 1, 1, 1), ["physout", "physin"] + VIRT)
            for _ in range(3)] for _ in range(2)]
    out = sq.apply_pepo_to_peps(a, sq.OPEPO(eye))
    for i in range(2):
        for j in range(3):
            t, r = out[i, j], a[i, j]
            np.testing.assert_allclose(
                np.transpose(t.data, [t.labels.index(x) for x in r.labels]), r.data, atol=0)
    # a PEPO of single-site Z operators: <a|Z...Z|a> against the dense brute force
    Z = np.diag([1.0, -1.0])
    zz = sq.OPEPO([[o.OTensor(Z.reshape(2, 2, 1, 1, 1, 1), ["physout", "physin"] + VIRT)
                    for _ in range(3)] for _ in range(2)])
    za = sq.apply_pepo_to_peps(a, zz)
    got = float(np.squeeze(sq.inner_product_peps(za, a).data))
    ids, n = bond_ids(2, 3)
    ops, p = [], 2 * n
    for i in range(2):
        for j in range(3):
            t = a[i, j]
            ops += [t.data, [p if lab == "phys" else ids[(i, j, lab)] for lab in t.labels]]
            ops += [Z, [p, p + 1]]
            ops += [t.data, [p + 1 if lab == "phys" else (ids[(i, j, lab)] + n if ids[(i, j, lab)] else 0)
                             for lab in t.labels]]
            p += 2
    np.testing.assert_allclose(got, np.einsum(*ops, [], optimize=True), rtol=1e-12)


def test_fliplr_and_can_contract_and_columns(rng):
    tn = square(random_grid(3, 4, 2, rng))
    m = tn.fliplr()
    assert (m.left_label, m.right_label) == ("right", "left")
    np.testing.assert_allclose(np.squeeze(m.exact_contract().data), np.squeeze(tn.exact_contract().data),
                               rtol=1e-12)
    assert tn.can_contract()
    assert not tn.is_left_right_periodic()
    bad = square(random_grid(3, 4, 2, rng))
    bad[1, 2] = o.OTensor(rng.standard_normal((3, 2, 2, 2)), VIRT)
    assert not bad.can_contract()
    col = tn.col_to_1D_array(1)
    assert (col.left_label, col.right_label) == ("up", "down") and len(col) == 3
    first = sq.column_to_mpo(tn, 0)
    assert isinstance(first, o.OMPS) and first.phys_label == "right"
    assert "left" not in first[0].labels
    last = sq.column_to_mpo(tn, 3)
    assert isinstance(last, o.OMPS) and last.phys_label == "left"
    mid = sq.column_to_mpo(tn, 1)
    assert isinstance(mid, o.OMPO) and (mid.physout_label, mid.physin_label) == ("right", "left")
