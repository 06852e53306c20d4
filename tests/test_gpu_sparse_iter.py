"""sparse(opr), Matrix, element access and the row/column iterators on the
device vs the oracle (GPU).  Sparsity pattern (colptr, rowval) bit-exact,
values within 1e-12 (absolute, entries are O(1))."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _heis(ed, hs, bonds, w=None):
    return ed.simplify(sum((ed.heisenberg_bond(hs, i, j, 1.0 if w is None else w[k]) for k, (i, j) in enumerate(bonds)), ed.NullOperator()))


def _reduced(ed, oracle, hs, n, irrep, H):
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    rh = ed.symmetry_reduce(hsr, irrep)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in irrep.symops_and_amplitudes()])
    red = oracle.symmetry_reduce(oracle.sector_basis(n, 0), n, perms, [a for _, a in irrep.symops_and_amplitudes()])
    return ed.represent(rh, H), oracle.Opr(red, ed.terms_arrays(H))


def _check_sparse(opr, o, tol=None):
    kw = {} if tol is None else {"tol": tol}
    S = opr.sparse(**kw)
    colptr, rowval, nz = o.sparse(**kw)
    assert np.array_equal(S.indptr, colptr)
    assert np.array_equal(S.indices, rowval)
    assert np.abs(S.data - nz).max(initial=0) < 1e-12


def test_sparse_sigma_plus_known(ed, golden):
    hs, s = ed.spin_half_system(4)
    opr = ed.represent(ed.represent(hs), s(1, "+"))
    dense = np.array(golden["known_answers"]["sigma_plus_site2_n4"]["dense"])
    assert np.array_equal(opr.sparse().toarray(), dense)
    assert np.array_equal(opr.Matrix(), dense)
    for i in range(16):
        assert np.array_equal(opr.get_row(i).toarray().ravel(), dense[i])
        assert np.array_equal(opr.get_column(i).toarray().ravel(), dense[:, i])
        for j in range(16):
            assert opr.get_element(i, j) == dense[i, j]
            assert opr[i, j] == dense[i, j]
    for bad in (-1, 16):
        with pytest.raises(ed.BoundsError):
            opr.get_row_iterator(bad)
        with pytest.raises(ed.BoundsError):
            opr.get_element(bad, 0)
        with pytest.raises(ed.BoundsError):
            opr[0, bad]


@pytest.mark.parametrize("k", [0, 2, 5])
def test_sparse_reduced_chain(ed, oracle, k):
    L = 12
    hs, _ = ed.spin_half_system(L)
    H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)] + [(i, (i + 2) % L) for i in range(L)], [1.0] * L + [0.3] * L)
    opr, o = _reduced(ed, oracle, hs, L, ed.TranslationIrrep(L, k), H)
    _check_sparse(opr, o)
    _check_sparse(opr, o, tol=0.0)


def test_sparse_non_hermitian(ed, oracle):
    """A non-Hermitian operator: the column iterator differs from the row one."""
    L = 10
    hs, s = ed.spin_half_system(L)
    op = ed.simplify(sum(((0.7 + 0.2j) * s(i, "+") * s((i + 1) % L, "-") + 0.3 * s(i, "z") for i in range(L)), ed.NullOperator()))
    opr, o = _reduced(ed, oracle, hs, L, ed.TranslationIrrep(L, 3), op)
    _check_sparse(opr, o)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(opr.dimension) + 1j * rng.standard_normal(opr.dimension)
    yt = x @ opr
    assert np.linalg.norm(yt - o.apply(np.zeros_like(x), x, transpose=True)) < 1e-12 * np.linalg.norm(yt)


@pytest.mark.parametrize("name,shape,k", [("chain24_k3", (24, 1), (3, 0)), ("torus4x6_pi0", (4, 6), (2, 0))])
def test_sparse_full_size_configs(ed, oracle, name, shape, k):
    n = shape[0] * shape[1]
    hs, _ = ed.spin_half_system(n)
    if shape[1] == 1:
        H = _heis(ed, hs, [(i, (i + 1) % n) for i in range(n)])
    else:
        j1, j2 = ed.square_bonds(*shape)
        H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    opr, o = _reduced(ed, oracle, hs, n, ed.TranslationIrrep(shape, k), H)
    _check_sparse(opr, o)


def test_iterators_vs_oracle(ed, oracle):
    L = 10
    hs, s = ed.spin_half_system(L)
    op = ed.simplify(sum((s(i, "+") * s((i + 1) % L, "-") * (1 + 0.1 * i) + s(i, "x") for i in range(L)), ed.NullOperator()))
    opr, o = _reduced(ed, oracle, hs, L, ed.TranslationIrrep(L, 2), op)
    for r in range(0, opr.dimension, 3):
        got = opr.get_row_iterator(r)
        idx, val = o.row_iterator(r)
        assert [g[0] for g in got] == idx.tolist()
        assert np.allclose([g[1] for g in got], val, atol=1e-14)
        got = opr.get_column_iterator(r)
        idx, val = o.col_iterator(r)
        assert [g[0] for g in got] == idx.tolist()
        assert np.allclose([g[1] for g in got], val, atol=1e-14)


def test_iterator_invalid_known(ed, golden):
    ka = golden["known_answers"]["iterator_invalid"]
    hs, _ = ed.spin_half_system(4)
    hsr = ed.represent(hs, ka["basis"])
    op = ed.SumOperator([ed.PureOperator(int(m), int(r), int(c), a) for m, r, c, a in ka["terms"]])
    opr = ed.represent(hsr, op)
    for r, expect in ka["row_iterator"].items():
        assert [list(p) for p in opr.get_row_iterator(int(r))] == expect
    for c, expect in ka["column_iterator"].items():
        assert [list(p) for p in opr.get_column_iterator(int(c))] == expect


def test_reduced_2x2_known(ed, golden):
    """test_reduced_representation.jl:159-231 (Pauli units)."""
    hs, s = ed.spin_half_system(4)
    j1 = sum((s(i, j) * s((i + 1) % 4, j) for i in range(4) for j in "xyz"), ed.NullOperator())
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    rh = ed.symmetry_reduce(hsr, ed.TranslationIrrep(4, 0))
    assert rh.dimension == 2 and rh.basis_list.tolist() == [0b0011, 0b0101]
    opr = ed.represent(rh, j1)  # unsimplified, as in the reference test
    H2 = np.array(golden["known_answers"]["reduced_2x2_n4_k0"]["matrix"])
    assert np.allclose(opr.Matrix(), H2, atol=1.5e-8)
    for r in range(2):
        row = np.zeros(2, complex)
        err = 0.0
        for i, v in opr.get_row_iterator(r):
            if 0 <= i < 2:
                row[i] += v
            else:
                err += v
        assert abs(err) < 1.5e-8
        assert np.allclose(row, H2[r], atol=1.5e-8)
    for i in range(2):
        for j in range(2):
            assert abs(opr.get_element(i, j) - H2[i, j]) < 1.5e-8


def test_spectrum_union_equals_full(ed):
    """test_reduced_representation.jl:236-252 on the device: the union over k of
    the reduced spectra equals the Sz = 0 spectrum (L = 10)."""
    L = 10
    hs, _ = ed.spin_half_system(L)
    H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)])
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    full = np.linalg.eigvalsh(ed.represent(hsr, H).Matrix())
    ev = []
    for k in range(L):
        M = ed.represent(ed.symmetry_reduce(hsr, ed.TranslationIrrep(L, k)), H).Matrix()
        assert np.allclose(M, M.conj().T, atol=1e-13)
        ev.extend(np.linalg.eigvalsh(M))
    assert np.allclose(np.sort(ev), full, atol=1e-10)


def _merged(pairs, n, tol):
    """Host merge of iterator pairs: valid indices, duplicates summed, |v| < tol dropped."""
    acc = {}
    for i, v in pairs:
        if 0 <= i < n:
            acc[i] = acc.get(i, 0) + v
    return {i: v for i, v in acc.items() if abs(v) >= tol}


@pytest.mark.parametrize("k", [0, 3])
def test_rows_columns_elements_vs_oracle(ed, oracle, k):
    """get_row / get_column / get_element / Matrix from the device walks over
    selected rows and columns against the oracle's dense matrix, on a
    non-Hermitian reduced operator (rows and columns differ)."""
    L = 10
    hs, s = ed.spin_half_system(L)
    op = ed.simplify(sum(((0.7 + 0.2j) * s(i, "+") * s((i + 1) % L, "-") + 0.3 * s(i, "z") * s((i + 2) % L, "z")
                          for i in range(L)), ed.NullOperator()))
    opr, o = _reduced(ed, oracle, hs, L, ed.TranslationIrrep(L, k), op)
    D = o.dense()
    n = opr.dimension
    M = opr.Matrix()
    assert np.abs(M - D).max() < 1e-14
    tol = np.sqrt(np.finfo(float).eps)
    for r in range(0, n, 5):
        row = opr.get_row(r)
        assert row.shape == (1, n)
        dr = D[r].copy()
        dr[np.abs(dr) < tol] = 0
        assert np.abs(row.toarray().ravel() - dr).max() < 1e-14
        assert np.array_equal(np.sort(row.indices), np.nonzero(dr)[0])
        col = opr.get_column(r)
        assert col.shape == (n, 1)
        dc = D[:, r].copy()
        dc[np.abs(dc) < tol] = 0
        assert np.abs(col.toarray().ravel() - dc).max() < 1e-14
        for c in range(0, n, 7):
            assert abs(opr.get_element(r, c) - o.get_element(r, c)) < 1e-14
            assert abs(opr[r, c] - D[r, c]) < 1e-14
    with pytest.raises(ed.BoundsError):
        opr.get_row(n)
    with pytest.raises(ed.BoundsError):
        opr.get_column(-1)
    with pytest.raises(ed.BoundsError):
        opr.get_element(0, n)


def test_get_row_6x6_without_sparse(ed):
    """A row and a column of the 6x6 J1-J2 k=(0,0) operator (252,091,362
    representatives) straight from the device walk: equal to the host merge of
    the row iterator, and row = column (a real symmetric matrix)."""
    hs, _ = ed.spin_half_system(36)
    j1, j2 = ed.square_bonds(6, 6)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    r = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ed.TranslationIrrep((6, 6), (0, 0)))
    opr = ed.represent(r, H)
    tol = np.sqrt(np.finfo(float).eps)
    for irow in (0, 12345, 98765432, r.dimension - 1):
        row = opr.get_row(irow)
        ref = _merged(opr.get_row_iterator(irow), r.dimension, tol)
        assert sorted(ref) == row.indices.tolist()
        got = dict(zip(row.indices.tolist(), row.data.tolist()))
        assert max(abs(got[i] - ref[i]) for i in ref) < 1e-13
        col = opr.get_column(irow)
        assert col.indices.tolist() == row.indices.tolist()
        assert np.abs(col.data - row.data).max() < 1e-13
        assert abs(opr.get_element(irow, irow) - ref.get(irow, 0.0)) < 1e-13
