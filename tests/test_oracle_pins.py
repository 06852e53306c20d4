"""Pins of the CPU oracle (no GPU).

Every oracle function is checked against known answers from the reference's
own tests (tests/golden/reference_known_answers.json), closed forms and brute
force, chosen so that a dropped term, a wrong sign or index, or a transposed
operand fails one of them.
"""
import math
from itertools import combinations

import numpy as np
import pytest

from conftest import (
    chain_bonds,
    dense_from_terms,
    heisenberg_terms_pauli,
    translation_chars,
    translation_group_perms,
)


# ---------------------------------------------------------------- basis (A1)
def test_basis_n4_known(oracle, golden):
    ka = golden["known_answers"]["hsr_n4_sz0_basis"]
    assert oracle.sector_basis(4, 0).tolist() == ka["basis"]


def test_basis_tj_known(oracle, golden):
    ka = golden["known_answers"]["tj_sectors"]
    site_qns = [[tuple(s[1]) for s in site] for site in ka["sites"]]
    assert [list(q) for q in oracle.quantum_number_sectors(site_qns)] == ka["sectors"]
    assert oracle.hs_get_basis_list(site_qns, set(map(tuple, ka["sectors"]))).tolist() == ka["full_basis"]
    for key, basis in ka["sector_basis"].items():
        q = tuple(int(x) for x in key.split(","))
        assert oracle.hs_get_basis_list(site_qns, {q}).tolist() == basis
    two = ka["two_sector_basis"]
    assert oracle.hs_get_basis_list(site_qns, set(map(tuple, two["sectors"]))).tolist() == two["basis"]
    assert oracle.hs_get_basis_list(site_qns, set()).tolist() == []


def test_basis_convention_n7_known(oracle, golden):
    ka = golden["known_answers"]["phase_convention_n7"]
    assert oracle.sector_basis(7, 5).tolist() == ka["plus5"]["basis"]
    assert oracle.sector_basis(7, -5).tolist() == ka["minus5"]["basis"]


@pytest.mark.parametrize("n", [1, 2, 5, 8, 11, 14])
def test_basis_brute_force(oracle, n):
    """Every reachable Sz sector equals the sorted words of that popcount."""
    words = np.arange(2 ** n, dtype=np.uint64)
    pc = np.array([bin(int(w)).count("1") for w in words])
    for p in range(n + 1):
        sz = n - 2 * p
        b = oracle.sector_basis(n, sz)
        assert np.array_equal(b, words[pc == p])
        assert b.size == math.comb(n, p)
    assert oracle.sector_basis(n, n + 1).size == 0  # unreachable sector


# ---------------------------------------------------------------- lookup (A2)
def test_lookup_semantics(oracle):
    keys = np.array([3, 5, 6, 9, 10, 12], np.uint64)
    q = np.array([0, 3, 4, 5, 6, 7, 9, 10, 11, 12, 13], np.uint64)
    expect = [-1, 0, -1, 1, 2, -1, 3, 4, -1, 5, -1]
    assert oracle.lookup(keys, q).tolist() == expect
    assert oracle.lookup(np.zeros(0, np.uint64), q[:2]).tolist() == [-1, -1]
    assert oracle.validate_sorted_unique(keys) == 0
    assert oracle.validate_sorted_unique(np.array([1, 3, 2], np.uint64)) == 1
    assert oracle.validate_sorted_unique(np.array([1, 2, 2], np.uint64)) == 2
    big = np.sort(np.random.default_rng(0).choice(2 ** 40, 5000, replace=False)).astype(np.uint64)
    qq = np.concatenate([big[::7], big[::11] + np.uint64(1)])
    ref = np.searchsorted(big, qq)
    ref = np.where((ref < big.size) & (big[np.minimum(ref, big.size - 1)] == qq), ref, -1)
    assert np.array_equal(oracle.lookup(big, qq), ref)


# ---------------------------------------------------------------- permutation (A4)
def test_symmetry_apply_known(oracle, golden):
    ka = golden["known_answers"]["symmetry_apply"]
    t = [x - 1 for x in ka["transop"]]
    for b, img in ka["bits"]:
        assert oracle.symmetry_apply(4, t, b) == img
    inv = [x - 1 for x in ka["invop"]]
    m, r, c, _ = ka["pop1"]
    assert [oracle.symmetry_apply(4, t, v) for v in (m, r, c)] == ka["pop1_trans"][:3]
    assert [oracle.symmetry_apply(4, inv, v) for v in (m, r, c)] == ka["pop1_inv"][:3]


def test_symmetry_apply_multibit_sites(oracle):
    # three 2-bit sites: fields move as blocks
    b = 0b10_01_11
    assert oracle.symmetry_apply(3, [1, 2, 0], b, bitwidths=[2, 2, 2]) == 0b01_11_10


# ---------------------------------------------------------------- reduce (A5)
def test_reduce_translation_n4_known(oracle, golden):
    ka = golden["known_answers"]["translation_n4"]
    parent = oracle.sector_basis(4, 0)
    perms = translation_group_perms(4)
    for idx, basis in ka["reduced_basis_by_irrep_index"].items():
        red = oracle.symmetry_reduce(parent, 4, perms, translation_chars(4, 1, int(idx) - 1, 0))
        assert red.basis_list.tolist() == basis


def test_reduce_point_and_space_group_known(oracle, golden):
    parent = oracle.sector_basis(4, 0)
    inv = np.array([0, 3, 2, 1], np.int32)
    ident = np.arange(4, dtype=np.int32)
    ka = golden["known_answers"]["inversion_n4"]["reduced_basis"]
    assert oracle.symmetry_reduce(parent, 4, np.stack([ident, inv]), [1, 1]).basis_list.tolist() == ka["even"]
    assert oracle.symmetry_reduce(parent, 4, np.stack([ident, inv]), [1, -1]).basis_list.tolist() == ka["odd"]
    # space group T x I: elements T^a and T^a I
    T = translation_group_perms(4)
    els, ch_even, ch_odd = [], [], []
    for a in range(4):
        els.append(T[a])
        ch_even.append(1)
        ch_odd.append(1)
    for a in range(4):
        els.append(T[a][inv])
        ch_even.append(1)
        ch_odd.append(-1)
    sg = golden["known_answers"]["space_group_n4"]["reduced_basis"]
    assert oracle.symmetry_reduce(parent, 4, np.stack(els), ch_even).basis_list.tolist() == sg["k0_even"]
    assert oracle.symmetry_reduce(parent, 4, np.stack(els), ch_odd).basis_list.tolist() == sg["k0_odd"]


@pytest.mark.parametrize("key,qn,sign", [("plus5", 5, 1), ("minus5", -5, -1)])
def test_reduce_phase_convention_n7(oracle, golden, key, qn, sign):
    ka = golden["known_answers"]["phase_convention_n7"]
    parent = oracle.sector_basis(7, qn)
    red = oracle.symmetry_reduce(parent, 7, translation_group_perms(7), translation_chars(7, 1, 1, 0))
    assert red.basis_list.tolist() == ka[key]["reduced"]
    psi = oracle.vector_unreduce(red, np.array([1.0 + 0j]))
    expect = np.array([np.exp(sign * 2j * np.pi * i / 7) / np.sqrt(7) for i in range(7)])
    assert np.allclose(psi, expect, atol=1.5e-8)


def test_reduce_dimensions_character_sum(oracle, golden):
    dims = golden["sector_dimensions"]
    parent = oracle.sector_basis(12, 0)
    perms = translation_group_perms(12)
    got = [oracle.symmetry_reduce(parent, 12, perms, translation_chars(12, 1, k, 0)).dimension for k in range(12)]
    assert got == dims["chain12_by_k"]
    parent = oracle.sector_basis(16, 0)
    perms = translation_group_perms(4, 4)
    for kx in range(4):
        for ky in range(4):
            d = oracle.symmetry_reduce(parent, 16, perms, translation_chars(4, 4, kx, ky)).dimension
            assert d == dims["torus4x4_by_k"][kx][ky]


def test_reduce_invariants(oracle):
    L = 10
    parent = oracle.sector_basis(L, 0)
    perms = translation_group_perms(L)
    total = 0
    rng = np.random.default_rng(5)
    for k in range(L):
        red = oracle.symmetry_reduce(parent, L, perms, translation_chars(L, 1, k, 0))
        total += red.dimension
        # mapping index: reps point to themselves; -1 exactly for excluded words
        for r, b in enumerate(red.basis_list):
            p = int(np.searchsorted(parent, b))
            assert red.basis_mapping_index[p] == r
        ex = red.basis_mapping_index < 0
        assert np.all(red.basis_mapping_amplitude[ex] == 0)
        counts = np.bincount(red.basis_mapping_index[~ex], minlength=red.dimension)
        assert np.allclose(np.abs(red.basis_mapping_amplitude[~ex]) ** 2, 1.0 / counts[red.basis_mapping_index[~ex]])
        assert np.all(L % counts == 0)  # orbit sizes divide |G|
        # orbit members share the representative's index and |U|^2 sums to 1 per rep
        w = np.zeros(red.dimension)
        np.add.at(w, red.basis_mapping_index[~ex], np.abs(red.basis_mapping_amplitude[~ex]) ** 2)
        assert np.allclose(w, 1.0)
        # reduce(unreduce(v)) == v
        v = rng.standard_normal(red.dimension) + 1j * rng.standard_normal(red.dimension)
        assert np.allclose(oracle.vector_reduce(red, oracle.vector_unreduce(red, v)), v, atol=1e-12)
    assert total == parent.size


def test_reduce_argument_errors(oracle):
    parent = oracle.sector_basis(4, 0)
    with pytest.raises(ValueError):
        oracle.symmetry_reduce(parent, 4, translation_group_perms(4), [1, 2, 1, 1])


# ---------------------------------------------------------------- elements (A7-A10)
def test_plain_matrix_known_n4(oracle, golden):
    parent = oracle.sector_basis(4, 0)
    o = oracle.Opr(parent, heisenberg_terms_pauli(chain_bonds(4)))
    assert np.array_equal(o.dense().real, np.array(golden["known_answers"]["heisenberg_n4_matrix_pauli"]["matrix"]))


def test_sigma_plus_sparse_known(oracle, golden):
    full = np.arange(16, dtype=np.uint64)
    # sigma^+ on the second site: |Up><Dn| = PureOperator(bit1, 0, bit1, 1)
    terms = (np.array([2], np.uint64), np.array([0], np.uint64), np.array([2], np.uint64), np.array([1.0 + 0j]))
    o = oracle.Opr(full, terms)
    S = o.scipy_sparse().toarray()
    assert np.array_equal(S.real.astype(int), np.array(golden["known_answers"]["sigma_plus_site2_n4"]["dense"]))
    # apply! adds, mul! (zero then apply) overwrites; row-vector form = x^T A
    x = np.arange(16) + 1j
    out = np.ones(16, complex)
    o.apply(out, x)
    assert np.allclose(out, 1 + S @ x)
    yt = o.apply(np.zeros(16, complex), x, transpose=True)
    assert np.allclose(yt, x @ S)


def test_iterator_invalid_known(oracle, golden):
    ka = golden["known_answers"]["iterator_invalid"]
    t = np.array(ka["terms"])
    terms = (t[:, 0].astype(np.uint64), t[:, 1].astype(np.uint64), t[:, 2].astype(np.uint64), t[:, 3].astype(complex))
    o = oracle.Opr(np.array(ka["basis"], np.uint64), terms)
    for r, expect in ka["row_iterator"].items():
        idx, val = o.row_iterator(int(r))
        assert [[int(i), v.real] for i, v in zip(idx, val)] == expect
    for c, expect in ka["column_iterator"].items():
        idx, val = o.col_iterator(int(c))
        assert [[int(i), v.real] for i, v in zip(idx, val)] == expect


def test_reduced_2x2_known(oracle, golden):
    parent = oracle.sector_basis(4, 0)
    red = oracle.symmetry_reduce(parent, 4, translation_group_perms(4), translation_chars(4, 1, 0, 0))
    o = oracle.Opr(red, heisenberg_terms_pauli(chain_bonds(4)))
    H2 = np.array(golden["known_answers"]["reduced_2x2_n4_k0"]["matrix"])
    assert np.allclose(o.dense(), H2, atol=1.5e-8)
    rows = np.zeros((2, 2), complex)
    for r in range(2):
        idx, val = o.row_iterator(r)
        for i, v in zip(idx, val):
            if i >= 0:
                rows[r, i] += v
    assert np.allclose(rows, H2, atol=1.5e-8)
    for i in range(2):
        for j in range(2):
            assert abs(o.get_element(i, j) - H2[i, j]) < 1.5e-8


def test_reduced_row_equals_column_n7(oracle):
    L = 7
    parent = oracle.sector_basis(L, 1)
    red = oracle.symmetry_reduce(parent, L, translation_group_perms(L), translation_chars(L, 1, 2, 0))
    o = oracle.Opr(red, heisenberg_terms_pauli(chain_bonds(L)))
    rows = np.zeros((o.dim, o.dim), complex)
    for r in range(o.dim):
        idx, val = o.row_iterator(r)
        for i, v in zip(idx, val):
            if i >= 0:
                rows[r, i] += v
    assert np.allclose(rows, o.dense(), atol=1e-8)
    assert np.allclose(o.dense(), o.dense().conj().T, atol=1e-12)


@pytest.mark.parametrize("L", [8, 10])
def test_spectrum_union_equals_full(oracle, L):
    """test_reduced_representation.jl:236-252: the union over k of the reduced
    spectra equals the Sz=0 spectrum (checked against a dense brute force)."""
    parent = oracle.sector_basis(L, 0)
    terms = heisenberg_terms_pauli(chain_bonds(L))
    full = np.linalg.eigvalsh(dense_from_terms(L, parent, terms))
    ev = []
    for k in range(L):
        red = oracle.symmetry_reduce(parent, L, translation_group_perms(L), translation_chars(L, 1, k, 0))
        ev.extend(np.linalg.eigvalsh(oracle.Opr(red, terms).dense()))
    assert np.allclose(np.sort(ev), full, atol=1e-9)


def test_oracle_matches_brute_force_dense(oracle):
    """Unreduced oracle matrix vs the straight PureOperator row rule."""
    L = 8
    parent = oracle.sector_basis(L, 0)
    terms = heisenberg_terms_pauli(chain_bonds(L), weights=np.linspace(0.5, 1.5, L))
    assert np.allclose(oracle.Opr(parent, terms).dense(), dense_from_terms(L, parent, terms))


# ---------------------------------------------------------------- sparse (A11)
def test_sparse_chop_after_merge(oracle):
    """choptol! runs after the duplicate merge (util.jl:88-94): two terms that
    cancel to below tol disappear, a sum just above tol stays."""
    basis = np.array([0, 1], np.uint64)
    eps = 1e-9
    terms = (
        np.array([1, 1, 1, 1], np.uint64),
        np.array([0, 0, 1, 1], np.uint64),
        np.array([1, 1, 0, 0], np.uint64),
        np.array([1.0, -1.0 + eps, 1e-7, 1e-7 + 0j]),
    )
    colptr, rowval, nz = oracle.Opr(basis, terms).sparse()
    # column 1 (word 1): entries from terms 0,1 at row 0 -> eps < tol -> chopped
    # column 0 (word 0): entries from terms 2,3 at row 1 -> 2e-7 > tol -> kept
    assert colptr.tolist() == [0, 1, 1]
    assert rowval.tolist() == [1]
    assert np.allclose(nz, [2e-7])


@pytest.mark.parametrize("reduced", [False, True])
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_sparse_parallel_equals_serial(oracle, reduced, threads):
    """sparse_parallel (abstract_operator_representation.jl:172-204): the
    column blocks built by each thread and concatenated in block order give
    exactly sparse_serial's CSC (same merge order within a column), for any
    thread count, including the chop-after-merge case."""
    L = 12
    parent = oracle.sector_basis(L, 0)
    terms = heisenberg_terms_pauli(chain_bonds(L) + [(i, (i + 3) % L) for i in range(L)],
                                   weights=np.linspace(0.5, 1.5, 2 * L))
    space = parent
    if reduced:
        space = oracle.symmetry_reduce(parent, L, translation_group_perms(L), translation_chars(L, 1, 5, 0))
    o = oracle.Opr(space, terms)
    a = o.sparse()
    b = o.sparse_parallel(n_threads=threads)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    basis = np.array([0, 1], np.uint64)
    chop = (np.array([1, 1, 1, 1], np.uint64), np.array([0, 0, 1, 1], np.uint64),
            np.array([1, 1, 0, 0], np.uint64), np.array([1.0, -1.0 + 1e-9, 1e-7, 1e-7 + 0j]))
    c = oracle.Opr(basis, chop).sparse_parallel(n_threads=threads)
    assert c[0].tolist() == [0, 1, 1] and c[1].tolist() == [1]


# ---------------------------------------------------------------- eigenvalues (A12)
def test_eigenvalues_brute_force(oracle, golden):
    ev = golden["eigenvalues_brute_force"]
    L = 12
    parent = oracle.sector_basis(L, 0)
    o = oracle.Opr(parent, heisenberg_terms_pauli(chain_bonds(L), weights=[0.25] * L))
    assert np.allclose(oracle.lowest_eigenvalues(o, 4), ev["chain12"], atol=1e-10)


@pytest.mark.parametrize("L", [8, 10])
def test_spectrum_union_with_spin_flip(oracle, L):
    """Translations x global spin flip (GlobalBitFlip, bitflipsymmetry.jl):
    the reduced dimensions over (k, parity) add up to C(L, L/2) and the union
    of the reduced spectra is the Sz=0 spectrum (dense brute force)."""
    parent = oracle.sector_basis(L, 0)
    terms = heisenberg_terms_pauli(chain_bonds(L))
    full = np.linalg.eigvalsh(dense_from_terms(L, parent, terms))
    perms = np.concatenate([translation_group_perms(L)] * 2)
    full_mask = (1 << L) - 1
    flips = np.array([0] * L + [full_mask] * L, np.uint64)
    ev, dims = [], 0
    for k in range(L):
        chi_t = translation_chars(L, 1, k, 0)
        for parity in (1, -1):
            chars = np.concatenate([chi_t, parity * chi_t])
            red = oracle.symmetry_reduce(parent, L, perms, chars, flips=flips)
            dims += red.dimension
            if red.dimension:
                ev.extend(np.linalg.eigvalsh(oracle.Opr(red, terms).dense()))
    assert dims == parent.size
    assert np.allclose(np.sort(ev), full, atol=1e-9)
