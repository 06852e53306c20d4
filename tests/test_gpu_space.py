"""Device enumeration, lookup and symmetry reduction vs the oracle (GPU)."""
import math

import numpy as np
import pytest

from conftest import translation_chars, translation_group_perms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,sz", [(4, 0), (7, 5), (7, -5), (12, 0), (16, 2), (20, 0)])
def test_sector_enumeration_bit_exact(ed, oracle, n, sz):
    hs, _ = ed.spin_half_system(n)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, sz))
    assert hsr.dimension == oracle.sector_basis(n, sz).size
    assert np.array_equal(hsr.basis_list, oracle.sector_basis(n, sz))


def test_full_space_and_multi_sector(ed, golden):
    hs, _ = ed.spin_half_system(4)
    assert ed.represent(hs).basis_list.tolist() == list(range(16))
    two = ed.represent(ed.HilbertSpaceSector(hs, [4, -2])).basis_list
    expect = sorted(w for w in range(16) if bin(w).count("1") in (0, 3))
    assert two.tolist() == expect
    assert ed.represent(ed.HilbertSpaceSector(hs, 1)).dimension == 0
    ka = golden["known_answers"]["hsr_n4_sz0_basis"]
    assert ed.represent(ed.HilbertSpaceSector(hs, 0)).basis_list.tolist() == ka["basis"]


def test_empty_space_operator(ed):
    """An operator on an empty sector: zero-length products, a 0 x 0 sparse
    and dense matrix (no kernel is launched on an empty grid)."""
    hs, _ = ed.spin_half_system(4)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 1))
    assert hsr.dimension == 0
    H = ed.simplify(sum((ed.heisenberg_bond(hs, i, (i + 1) % 4, 1.0) for i in range(4)), ed.NullOperator()))
    opr = ed.represent(hsr, H)
    assert opr.dimension == 0
    out = np.zeros(0)
    ed.mul_(out, opr, np.zeros(0))
    ed.mul_(out, np.zeros(0), opr)
    assert out.shape == (0,)
    assert opr.sparse().shape == (0, 0)
    assert opr.Matrix().shape == (0, 0)


def test_explicit_list_and_lookup(ed, oracle):
    hs, _ = ed.spin_half_system(20)
    rng = np.random.default_rng(11)
    words = np.unique(rng.integers(0, 2 ** 20, 5000, dtype=np.uint64))
    hsr = ed.represent(hs, rng.permutation(words))  # represent_array sorts its input
    assert np.array_equal(hsr.basis_list, words)
    q = np.concatenate([words[::3], rng.integers(0, 2 ** 20, 3000, dtype=np.uint64), [0, 2 ** 20 - 1, 2 ** 40]])
    q = q.astype(np.uint64)
    assert np.array_equal(hsr.lookup(q), oracle.lookup(words, q))
    sec = ed.represent(ed.HilbertSpaceSector(hs, 0))  # rank-indexed lookup of a full sector
    qs = np.concatenate([oracle.sector_basis(20, 0)[::97], q])
    assert np.array_equal(sec.lookup(qs), oracle.lookup(oracle.sector_basis(20, 0), qs))


def test_argument_errors(ed):
    hs9, _ = ed.spin_half_system(9)
    with pytest.raises(ed.ArgumentError):
        ed.represent(hs9, "UInt8")  # BR too narrow (hilbert_space_representation.jl:63-69)
    hs4, _ = ed.spin_half_system(4)
    hsr = ed.represent(ed.HilbertSpaceSector(hs4, 0))
    with pytest.raises(ed.ArgumentError):  # |chi| != 1 (symmetry_reduce_generic.jl:27-29)
        ed.symmetry_reduce(hsr, [(p, 2.0) for p in ed.translation_group(4)])
    with pytest.raises(ed.ArgumentError):  # identity must come first
        ed.symmetry_reduce(hsr, [(p, 1.0) for p in ed.translation_group(4)[::-1]])
    with pytest.raises(ed.ArgumentError):  # not a group
        ed.symmetry_reduce(hsr, [(p, 1.0) for p in ed.translation_group(4)[:3]])


def test_reduce_known_answers(ed, golden):
    hs, _ = ed.spin_half_system(4)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    for idx, basis in golden["known_answers"]["translation_n4"]["reduced_basis_by_irrep_index"].items():
        assert ed.symmetry_reduce(hsr, ed.TranslationIrrep(4, int(idx) - 1)).basis_list.tolist() == basis
    e, inv = ed.SitePermutation([1, 2, 3, 4]), ed.SitePermutation([1, 4, 3, 2])
    ka = golden["known_answers"]["inversion_n4"]["reduced_basis"]
    r = ed.symmetry_reduce(hsr, [(e, 1.0), (inv, 1.0)])
    assert r.basis_list.tolist() == ka["even"] and r.canonicaliser() == "generic"
    assert ed.symmetry_reduce(hsr, [(e, 1.0), (inv, -1.0)]).basis_list.tolist() == ka["odd"]
    T = ed.translation_group(4)
    sg = golden["known_answers"]["space_group_n4"]["reduced_basis"]
    els = T + [t * inv for t in T]
    assert ed.symmetry_reduce(hsr, [(g, 1.0) for g in els]).basis_list.tolist() == sg["k0_even"]
    odd = [(g, 1.0) for g in T] + [(t * inv, -1.0) for t in T]
    assert ed.symmetry_reduce(hsr, odd).basis_list.tolist() == sg["k0_odd"]


@pytest.mark.parametrize("qn,sign", [(5, 1), (-5, -1)])
def test_phase_convention_n7(ed, golden, qn, sign):
    """test_symmetry_reduce.jl:164-214 through the exported mapping U."""
    hs, _ = ed.spin_half_system(7)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, qn))
    r = ed.symmetry_reduce(hsr, ed.TranslationIrrep(7, 1))
    key = "plus5" if qn > 0 else "minus5"
    assert r.basis_list.tolist() == golden["known_answers"]["phase_convention_n7"][key]["reduced"]
    U = r.basis_mapping_amplitude
    expect = np.array([np.exp(sign * 2j * np.pi * i / 7) / np.sqrt(7) for i in range(7)])
    assert np.allclose(U, expect, atol=1e-15)
    assert r.basis_mapping_index.tolist() == [0] * 7


def _check_reduce(ed, oracle, hs, hsr, parent, ops, n):
    r = ed.symmetry_reduce(hsr, ops)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in ops])
    red = oracle.symmetry_reduce(parent, n, perms, [a for _, a in ops])
    assert np.array_equal(r.basis_list, red.basis_list)
    assert np.array_equal(r.basis_mapping_index, red.basis_mapping_index)
    assert np.abs(r.basis_mapping_amplitude - red.basis_mapping_amplitude).max(initial=0) < 1e-14
    return r, red


@pytest.mark.parametrize("k", range(12))
def test_reduce_chain12_all_k(ed, oracle, k):
    hs, _ = ed.spin_half_system(12)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    r, _ = _check_reduce(ed, oracle, hs, hsr, oracle.sector_basis(12, 0), ed.TranslationIrrep(12, k).symops_and_amplitudes(), 12)
    assert r.canonicaliser() == "torus"


def test_reduce_square4x4_all_k(ed, oracle, golden):
    hs, _ = ed.spin_half_system(16)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    parent = oracle.sector_basis(16, 0)
    for kx in range(4):
        for ky in range(4):
            r, _ = _check_reduce(ed, oracle, hs, hsr, parent, ed.TranslationIrrep((4, 4), (kx, ky)).symops_and_amplitudes(), 16)
            assert r.dimension == golden["sector_dimensions"]["torus4x4_by_k"][kx][ky]


@pytest.mark.parametrize("k,parity", [(0, 1), (0, -1), (8, 1), (8, -1)])
def test_reduce_dihedral_generic(ed, oracle, k, parity):
    """Translations x inversion of a 16-site ring (32 elements): the
    translations are all present, so the torus canonicaliser runs over the two
    cosets {T, T I}.  The 1D irreps have chi(T^a) = exp(-2 pi i k a / L) real
    (k = 0, L/2) and chi(T^a I) = parity * chi(T^a)."""
    L = 16
    hs, _ = ed.spin_half_system(L)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    T = ed.translation_group(L)
    inv = ed.SitePermutation([((-i) % L) + 1 for i in range(L)])
    irrep = ed.TranslationIrrep(L, k)
    ch = [irrep.character(a).real for a in range(L)]
    ops = [(t, ch[a]) for a, t in enumerate(T)] + [(t * inv, parity * ch[a]) for a, t in enumerate(T)]
    r, _ = _check_reduce(ed, oracle, hs, hsr, oracle.sector_basis(L, 0), ops, L)
    assert r.canonicaliser() == "torus"


@pytest.mark.parametrize("k,parity", [(0, 1), (0, -1), (4, 1), (4, -1)])
def test_reduce_even_dihedral_generic(ed, oracle, k, parity):
    """Even translations x inversion of a 16-site ring (16 elements, not all
    translations): the generic LDS-table canonicaliser; chi(T^2a) =
    exp(-2 pi i k 2a / L) is real for k = 0, 4."""
    L = 16
    hs, _ = ed.spin_half_system(L)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    T = ed.translation_group(L)
    inv = ed.SitePermutation([((-i) % L) + 1 for i in range(L)])
    irrep = ed.TranslationIrrep(L, k)
    ch = [irrep.character(a).real for a in range(L)]
    ops = [(T[a], ch[a]) for a in range(0, L, 2)] + [(T[a] * inv, parity * ch[a]) for a in range(0, L, 2)]
    r, _ = _check_reduce(ed, oracle, hs, hsr, oracle.sector_basis(L, 0), ops, L)
    assert r.canonicaliser() == "generic"


def test_reduce_multibit_sites_from_list(ed, oracle):
    """Sites of three states (2 bits each) given by an explicit basis list;
    cyclic site permutations act on 2-bit fields (generic canonicaliser)."""
    n = 6
    s3 = ed.Site([ed.State("a", 0), ed.State("b", 0), ed.State("c", 0)])
    hs = ed.HilbertSpace([s3] * n)
    words = []
    for code in range(3 ** n):
        w, c = 0, code
        for i in range(n):
            w |= (c % 3) << (2 * i)
            c //= 3
        words.append(w)
    words = np.array(sorted(words), np.uint64)
    hsr = ed.represent(hs, words)
    ops = [(ed.translation_perm(n, a), np.exp(-2j * np.pi * 2 * a / n)) for a in range(n)]
    r = ed.symmetry_reduce(hsr, ops)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in ops])
    assert r.canonicaliser() == "generic"
    red = oracle.symmetry_reduce(words, 2 * n, perms, [a for _, a in ops])
    assert np.array_equal(r.basis_list, red.basis_list)
    assert np.array_equal(r.basis_mapping_index, red.basis_mapping_index)
    assert np.abs(r.basis_mapping_amplitude - red.basis_mapping_amplitude).max() < 1e-14


@pytest.mark.parametrize("name,shape,k,n_up", [("chain24_k3", (24, 1), (3, 0), 12), ("torus4x6_pi0", (4, 6), (2, 0), 12)])
def test_reduce_full_size_configs(ed, oracle, golden, name, shape, k, n_up):
    """The 24-site configurations against the oracle's symmetry_reduce over
    the 2,704,156-state parent: basis, mapping index bit-exact."""
    n = shape[0] * shape[1]
    hs, _ = ed.spin_half_system(n)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    r, red = _check_reduce(ed, oracle, hs, hsr, oracle.sector_basis(n, 0), ed.TranslationIrrep(shape, k).symops_and_amplitudes(), n)
    assert r.dimension == golden["sector_dimensions"][name]


def test_reduce_6x6_properties(ed):
    """6x6, k=(0,0): closed-form dimension, ascending unique representatives,
    every representative is the minimum of its orbit, orbit sizes divide 36 and
    add up to C(36,18) (every orbit is compatible at k = 0)."""
    hs, _ = ed.spin_half_system(36)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    assert hsr.dimension == math.comb(36, 18)
    r = ed.symmetry_reduce(hsr, ed.TranslationIrrep((6, 6), (0, 0)))
    assert r.dimension == 252091362
    assert r.canonicaliser() == "torus"
    orb = r.orbit_sizes.astype(np.int64)
    assert np.all(36 % orb == 0)
    assert int(orb.sum()) == math.comb(36, 18)
    b = r.basis_list
    assert np.all(b[1:] > b[:-1])
    assert np.all(np.array([bin(int(x)).count("1") for x in b[:: 100003]]) == 18)
    # sample: every translate of a representative is >= it, and the number of
    # distinct translates is its orbit size
    rng = np.random.default_rng(0)
    sample = rng.choice(b.size, 3000, replace=False)
    full = (1 << 36) - 1
    for i in sample:
        w = int(b[i])
        imgs = set()
        for bb in range(6):
            for a in range(6):
                out = 0
                for s in range(36):
                    if (w >> s) & 1:
                        x, y = s % 6, s // 6
                        out |= 1 << (((x + a) % 6) + 6 * ((y + bb) % 6))
                imgs.add(out & full)
        assert min(imgs) == w
        assert len(imgs) == orb[i]


@pytest.mark.parametrize("k", [0, 5])
def test_ring63_translations_vs_oracle(ed, oracle, k):
    """The widest words a UInt64 space allows (63 sites, hilbert_space_representation.jl:63-69):
    a 63-site ring with two up spins reduced by its 63 translations at
    momentum k (complex characters at k = 5): basis, mapping bit-exact, and
    the XY products (row and column forms) against the oracle."""
    n = 63
    hs, s = ed.spin_half_system(n)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 4 - n))
    parent = oracle.sector_basis(n, 4 - n)
    assert np.array_equal(hsr.basis_list, parent)
    irrep = ed.TranslationIrrep(n, k)
    rh = ed.symmetry_reduce(hsr, irrep)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in irrep.symops_and_amplitudes()])
    red = oracle.symmetry_reduce(parent, n, perms, [a for _, a in irrep.symops_and_amplitudes()])
    assert np.array_equal(rh.basis_list, red.basis_list)
    assert np.array_equal(rh.basis_mapping_index, red.basis_mapping_index)
    assert np.abs(rh.basis_mapping_amplitude - red.basis_mapping_amplitude).max(initial=0) < 1e-14
    H = ed.simplify(sum((ed.heisenberg_bond(hs, i, (i + 1) % n, 1.0, 0.0) for i in range(n)), ed.NullOperator()))
    opr = ed.represent(rh, H)
    o = oracle.Opr(red, ed.terms_arrays(H))
    rng = np.random.default_rng(k)
    x = rng.standard_normal(rh.dimension) + 1j * rng.standard_normal(rh.dimension)
    for transpose in (False, True):
        ref = o.apply(np.zeros(rh.dimension, complex), x, transpose=transpose)
        out = np.zeros(rh.dimension, complex)
        if transpose:
            ed.mul_(out, x, opr)
        else:
            ed.mul_(out, opr, x)
        assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
