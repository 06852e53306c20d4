"""Translations (and other site permutations) times the global spin flip
(GlobalBitFlip, bitflipsymmetry.jl:8-38) on the device vs the oracle: reduced
basis and mapping bit-exact, products and eigenvalues (GPU)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _heis(ed, hs, bonds, w=None):
    return ed.simplify(sum((ed.heisenberg_bond(hs, i, j, 1.0 if w is None else w[k]) for k, (i, j) in enumerate(bonds)),
                           ed.NullOperator()))


def _flip_irrep(ed, base, parity):
    return ed.make_product_irrep([(ed.GlobalBitFlip(False), 1.0), (ed.GlobalBitFlip(True), float(parity))], base)


def _oracle_reduce(ed, oracle, hs, ops, n):
    perms, flips = [], []
    for op, _ in ops:
        p, f = ed.representation._perm_and_flip(hs, op)
        perms.append(p)
        flips.append(f)
    return oracle.symmetry_reduce(oracle.sector_basis(n, 0), n, np.stack(perms), [a for _, a in ops],
                                  flips=np.array(flips, np.uint64))


def _system(ed, oracle, shape, k, parity, generic=False):
    Lx, Ly = shape
    n = Lx * Ly
    hs, _ = ed.spin_half_system(n)
    if Ly == 1:
        bonds = [(i, (i + 1) % n) for i in range(n)]
        w = None
    else:
        j1, j2 = ed.square_bonds(Lx, Ly)
        bonds = j1 + j2
        w = [1.0] * len(j1) + [0.5] * len(j2)
    H = _heis(ed, hs, bonds, w)
    if generic == "dihedral":
        # translations x inversion x flip of a ring: 4L elements, 4 cosets T p
        T = ed.translation_group(n)
        inv = ed.SitePermutation([((-i) % n) + 1 for i in range(n)])
        ch = [ed.TranslationIrrep(n, k[0]).character(a).real for a in range(n)]
        base = [(t, ch[a]) for a, t in enumerate(T)] + [(t * inv, ch[a]) for a, t in enumerate(T)]
    elif generic:
        # even translations x inversion x flip: 2L elements, generic tables with XOR
        T = ed.translation_group(n)
        inv = ed.SitePermutation([((-i) % n) + 1 for i in range(n)])
        ch = [ed.TranslationIrrep(n, k[0]).character(a).real for a in range(n)]
        base = [(T[a], ch[a]) for a in range(0, n, 2)] + [(T[a] * inv, ch[a]) for a in range(0, n, 2)]
    else:
        base = ed.TranslationIrrep(shape if Ly > 1 else n, k if Ly > 1 else k[0])
    ops = _flip_irrep(ed, base, parity)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    r = ed.symmetry_reduce(hsr, ops)
    red = _oracle_reduce(ed, oracle, hs, ops, n)
    return hs, H, r, red


CASES = [((12, 1), (0, 0), 1, False), ((12, 1), (3, 0), -1, False), ((12, 1), (6, 0), 1, False),
         ((4, 4), (0, 0), -1, False), ((4, 4), (1, 2), 1, False), ((4, 4), (2, 0), -1, False),
         ((6, 4), (3, 0), 1, False), ((3, 4), (1, 1), -1, False),
         ((12, 1), (0, 0), 1, "dihedral"), ((12, 1), (6, 0), -1, "dihedral"),
         ((12, 1), (0, 0), 1, True), ((12, 1), (3, 0), -1, True)]


@pytest.mark.parametrize("shape,k,parity,generic", CASES)
def test_reduce_and_apply_vs_oracle(ed, oracle, shape, k, parity, generic):
    hs, H, r, red = _system(ed, oracle, shape, k, parity, generic)
    assert r.canonicaliser() == ("generic" if generic is True else "torus")
    assert np.array_equal(r.basis_list, red.basis_list)
    assert np.array_equal(r.basis_mapping_index, red.basis_mapping_index)
    assert np.abs(r.basis_mapping_amplitude - red.basis_mapping_amplitude).max(initial=0) < 1e-14
    if r.dimension == 0:
        return
    opr = ed.represent(r, H)
    o = oracle.Opr(red, ed.terms_arrays(H))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(r.dimension) + 1j * rng.standard_normal(r.dimension)
    for transpose in (False, True):
        ref = o.apply(np.zeros(r.dimension, complex), x, transpose=transpose)
        out = np.zeros(r.dimension, complex)
        if transpose:
            ed.mul_(out, x, opr)
        else:
            ed.mul_(out, opr, x)
        assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
    # vector maps
    v = rng.standard_normal(red.parent.size) + 1j * rng.standard_normal(red.parent.size)
    assert np.abs(ed.symmetry_reduce(r, v) - oracle.vector_reduce(red, v)).max() < 1e-12
    small = ed.symmetry_reduce(r, v)
    assert np.abs(ed.symmetry_unreduce(r, small) - oracle.vector_unreduce(red, small)).max() < 1e-12


def test_chain16_ground_state_over_k_and_parity(ed, golden):
    """The lowest eigenvalue over the 32 (k, parity) sectors of the 16-site
    ring is the Sz=0 ground state."""
    n = 16
    hs, _ = ed.spin_half_system(n)
    H = _heis(ed, hs, [(i, (i + 1) % n) for i in range(n)])
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    dims, e0 = 0, np.inf
    for k in range(n):
        for parity in (1, -1):
            r = ed.symmetry_reduce(hsr, _flip_irrep(ed, ed.TranslationIrrep(n, k), parity))
            dims += r.dimension
            if r.dimension:
                e0 = min(e0, ed.eigs(ed.represent(r, H), 1, tol=1e-12, max_basis=min(120, r.dimension)).values[0])
    assert dims == math.comb(16, 8)
    ref = golden["eigenvalues_brute_force"]["chain16"][0]
    assert abs(e0 - ref) < 1e-11 * abs(ref)


def test_torus6x6_parity_sectors(ed):
    """6x6 J1-J2, k = (0,0): the even and odd spin-flip sectors split the
    252,091,362 representatives; the even sector holds the ground state of the
    k = (0,0) sector (E0 = -18.137147540, the unreduced-flip Lanczos value)."""
    hs, _ = ed.spin_half_system(36)
    j1, j2 = ed.square_bonds(6, 6)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    dims = {}
    for parity in (1, -1):
        r = ed.symmetry_reduce(hsr, _flip_irrep(ed, ed.TranslationIrrep((6, 6), (0, 0)), parity))
        assert r.canonicaliser() == "torus"
        dims[parity] = r.dimension
        if parity == 1:
            opr = ed.represent(r, H)
            e = ed.eigs(opr, 1, tol=1e-10, max_basis=100, criterion="eigenvalue").values[0]
            assert abs(e - (-18.137147540)) < 1e-8
    assert dims[1] + dims[-1] == 252091362


@pytest.mark.parametrize("k", [(0, 0), (2, 2)])
@pytest.mark.parametrize("point", ["A1", "A2", "B1", "B2"])
@pytest.mark.parametrize("parity", [0, 1, -1])
def test_space_group_4x4_vs_oracle(ed, oracle, k, point, parity):
    """Translations x C4v (128 elements), optionally x spin flip (256): the
    torus canonicaliser over cosets T p; basis, mapping and products vs the
    oracle's generic reduction."""
    n = 16
    hs, _ = ed.spin_half_system(n)
    j1, j2 = ed.square_bonds(4, 4)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    base = ed.SquareSpaceIrrep(4, k, point)
    ops = base.symops_and_amplitudes() if parity == 0 else _flip_irrep(ed, base, parity)
    r = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ops)
    red = _oracle_reduce(ed, oracle, hs, ops, n)
    assert r.canonicaliser() == "torus"
    assert np.array_equal(r.basis_list, red.basis_list)
    assert np.array_equal(r.basis_mapping_index, red.basis_mapping_index)
    assert np.abs(r.basis_mapping_amplitude - red.basis_mapping_amplitude).max(initial=0) < 1e-14
    if r.dimension == 0:
        return
    opr = ed.represent(r, H)
    o = oracle.Opr(red, ed.terms_arrays(H))
    x = np.random.default_rng(1).standard_normal(r.dimension)
    ref = o.apply(np.zeros(r.dimension, complex), x.astype(complex))
    out = np.zeros(r.dimension)
    ed.mul_(out, opr, x)
    assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
    E = ed.eigs(opr, 1, tol=1e-12, max_basis=min(100, r.dimension)).values[0]
    assert abs(E - oracle.lowest_eigenvalues(o, 1)[0]) < 1e-11 * abs(E)


@pytest.mark.parametrize("k,names", [((2, 0), ["A1", "A2", "B1", "B2"]), ((1, 0), ["A'", "A''"]),
                                     ((1, 3), ["A'", "A''"])])
def test_little_groups_4x4_vs_oracle(ed, oracle, k, names):
    """Little groups away from (0,0) and (pi,pi) (symmetry_reduce_symmorphic.jl
    keeps the point elements of nonzero character): C2v at (pi, 0) and Cs on
    the axial and anti-diagonal mirror lines, translations x little group
    (64 or 32 elements, complex characters at k = (1, 0) and (1, 3)); basis,
    mapping and products vs the oracle's generic reduction, and the lowest
    eigenvalue over the irreps equal to the translation sector's."""
    n = 16
    hs, _ = ed.spin_half_system(n)
    j1, j2 = ed.square_bonds(4, 4)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    e_irreps = []
    for name in names:
        ops = ed.SquareSpaceIrrep(4, k, name).symops_and_amplitudes()
        r = ed.symmetry_reduce(hsr, ops)
        red = _oracle_reduce(ed, oracle, hs, ops, n)
        assert np.array_equal(r.basis_list, red.basis_list)
        assert np.array_equal(r.basis_mapping_index, red.basis_mapping_index)
        assert np.abs(r.basis_mapping_amplitude - red.basis_mapping_amplitude).max(initial=0) < 1e-14
        opr = ed.represent(r, H)
        o = oracle.Opr(red, ed.terms_arrays(H))
        rng = np.random.default_rng(2)
        x = rng.standard_normal(r.dimension) + 1j * rng.standard_normal(r.dimension)
        for transpose in (False, True):
            ref = o.apply(np.zeros(r.dimension, complex), x, transpose=transpose)
            out = np.zeros(r.dimension, complex)
            if transpose:
                ed.mul_(out, x, opr)
            else:
                ed.mul_(out, opr, x)
            assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref), (name, transpose)
        e_irreps.append(ed.eigs(opr, 1, tol=1e-12, max_basis=min(100, r.dimension)).values[0])
    rk = ed.symmetry_reduce(hsr, ed.TranslationIrrep((4, 4), k))
    e_k = ed.eigs(ed.represent(rk, H), 1, tol=1e-12, max_basis=100).values[0]
    assert abs(min(e_irreps) - e_k) < 1e-10 * abs(e_k)


@pytest.mark.parametrize("flip", [False, True])
def test_space_group_6x6_a1(ed, flip):
    """6x6 J1-J2, k = (0,0), point irrep A1 of C4v (288 elements; 576 with the
    even spin flip): the ground state of the k = (0,0) translation sector lies
    in it (E0 = -18.137147540, the translation-only value)."""
    hs, _ = ed.spin_half_system(36)
    j1, j2 = ed.square_bonds(6, 6)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    base = ed.SquareSpaceIrrep(6, (0, 0), "A1")
    ops = _flip_irrep(ed, base, 1) if flip else base
    r = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ops)
    assert r.canonicaliser() == "torus"
    assert r.dimension == (15804956 if flip else 31566122)
    orb = r.orbit_sizes.astype(np.int64)
    assert np.all((576 if flip else 288) % orb == 0)
    e = ed.eigs(ed.represent(r, H), 1, tol=1e-10, max_basis=100, criterion="eigenvalue", mode="mixed").values[0]
    assert abs(e - (-18.137147540)) < 1e-8
