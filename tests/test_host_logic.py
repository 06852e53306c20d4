"""Host-side logic of the product (no GPU): the operator term format and its
algebra, Hilbert-space sectors, symmetry groups, and the C-ABI library."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, dense_from_terms


def _dense_of(ed, n, op):
    words = np.arange(2 ** n, dtype=np.uint64)
    return dense_from_terms(n, words, ed.terms_arrays(op))


def _kron_site(n, i, m):
    out = np.array([[1.0 + 0j]])
    for j in reversed(range(n)):
        out = np.kron(out, m if j == i else np.eye(2))
    return out


SX = np.array([[0, 1], [1, 0]], complex)
SY = np.array([[0, -1j], [1j, 0]], complex)
SZ = np.array([[1, 0], [0, -1]], complex)
SP = np.array([[0, 1], [0, 0]], complex)
SM = np.array([[0, 0], [1, 0]], complex)


@pytest.mark.parametrize("j,m", [("x", SX), ("y", SY), ("z", SZ), ("+", SP), ("-", SM)])
def test_pauli_matrix_vs_kron(ed, j, m):
    """pauli_matrix (spin_half.jl:28-42) with Up = state 0 = bit 0 clear."""
    hs, _ = ed.spin_half_system(3)
    for i in range(3):
        assert np.allclose(_dense_of(ed, 3, ed.pauli_matrix(hs, i, j)), _kron_site(3, i, m))


def test_products_and_sums_vs_kron(ed):
    hs, s = ed.spin_half_system(3)
    op = s(0, "x") * s(1, "y") + 2.5 * s(2, "z") - s(0, "+") * s(0, "-")
    ref = _kron_site(3, 0, SX) @ _kron_site(3, 1, SY) + 2.5 * _kron_site(3, 2, SZ) - _kron_site(3, 0, SP @ SM)
    assert np.allclose(_dense_of(ed, 3, op), ref)
    assert np.allclose(_dense_of(ed, 3, ed.simplify(op)), ref)


def test_pure_operator_validation(ed):
    with pytest.raises(ed.ArgumentError):
        ed.PureOperator(0b01, 0b10, 0b00, 1.0)
    with pytest.raises(ed.ArgumentError):
        ed.PureOperator(0b01, 0b00, 0b10, 1.0)
    # product rule: pure_operator.jl:136-154
    a = ed.PureOperator(0b011, 0b001, 0b010, 2.0)
    b = ed.PureOperator(0b110, 0b010, 0b100, 3.0)
    assert a * b == ed.PureOperator(0b111, 0b001, 0b100, 6.0)
    c = ed.PureOperator(0b010, 0b000, 0b000, 1.0)
    assert isinstance(a * c, ed.NullOperator)


def test_simplify_heisenberg_known(ed):
    """test_operator_simplify.jl:68-85: sum of sigma.sigma simplifies to the
    term set of 2 s+s- + 2 s-s+ + szsz."""
    n = 4
    hs, s = ed.spin_half_system(n)
    h1 = sum((2 * s(i, "+") * s((i + 1) % n, "-") for i in range(n)), ed.NullOperator())
    h1 = h1 + sum((2 * s(i, "-") * s((i + 1) % n, "+") for i in range(n)), ed.NullOperator())
    h1 = h1 + sum((s(i, "z") * s((i + 1) % n, "z") for i in range(n)), ed.NullOperator())
    h2 = sum((s(i, j) * s((i + 1) % n, j) for i in range(n) for j in "xyz"), ed.NullOperator())
    h3, h4 = ed.simplify(h1), ed.simplify(h2)
    assert len(h1.terms) < len(h2.terms)
    assert {(t.bitmask, t.bitrow, t.bitcol, complex(t.amplitude)) for t in h3.terms} == {
        (t.bitmask, t.bitrow, t.bitcol, complex(t.amplitude)) for t in h4.terms
    }
    assert all(not isinstance(t.amplitude, complex) for t in h4.terms)  # demoted to real
    assert len(h4.terms) == 6 * n  # the |00><11| parts cancel
    keys = [(t.bitmask, t.bitrow, t.bitcol) for t in h4.terms]
    assert keys == sorted(keys)
    assert isinstance(ed.simplify(h3 - h3), ed.NullOperator)


def test_heisenberg_bond_units(ed):
    hs, _ = ed.spin_half_system(2)
    bond = ed.heisenberg_bond(hs, 0, 1, J=1.0, Delta=0.5)
    ref = 0.25 * (_kron_site(2, 0, SX) @ _kron_site(2, 1, SX) + _kron_site(2, 0, SY) @ _kron_site(2, 1, SY))
    ref += 0.25 * 0.5 * _kron_site(2, 0, SZ) @ _kron_site(2, 1, SZ)
    assert np.allclose(_dense_of(ed, 2, bond), ref)


def test_sectors_and_popcounts(ed):
    from edhip.hilbert import popcount_sets

    hs, _ = ed.spin_half_system(4)
    assert popcount_sets(ed.HilbertSpaceSector(hs, 0)) == [2]
    assert popcount_sets(ed.HilbertSpaceSector(hs, [4, -2])) == [0, 3]
    assert popcount_sets(ed.HilbertSpaceSector(hs, 1)) == []  # unreachable
    assert popcount_sets(hs) == [0, 1, 2, 3, 4]
    n7, _ = ed.spin_half_system(7)
    assert popcount_sets(ed.HilbertSpaceSector(n7, 5)) == [1]
    assert popcount_sets(ed.HilbertSpaceSector(n7, -5)) == [6]
    up, dn, em = ed.State("Up", (1, 1)), ed.State("Dn", (1, -1)), ed.State("Em", (0, 0))
    tj = ed.HilbertSpace([ed.Site([em, up, dn]), ed.Site([up, dn]), ed.Site([up, dn])])
    assert popcount_sets(tj) is None  # general sites: explicit basis lists
    assert tj.bitwidths == [2, 1, 1] and tj.bitoffsets == [0, 2, 3, 4]


def test_symmetry_apply_host_known(ed, golden):
    ka = golden["known_answers"]["symmetry_apply"]
    hs, s = ed.spin_half_system(4)
    T = ed.SitePermutation(ka["transop"])
    I = ed.SitePermutation(ka["invop"])
    for b, img in ka["bits"]:
        assert ed.symmetry_apply(hs, T, b) == (img, 1)
    p1 = ed.PureOperator(*ka["pop1"])
    assert ed.symmetry_apply(hs, T, p1) == ed.PureOperator(*ka["pop1_trans"])
    assert ed.symmetry_apply(hs, I, p1) == ed.PureOperator(*ka["pop1_inv"])
    j1 = ed.simplify(sum((s(i, j) * s((i + 1) % 4, j) for i in range(4) for j in "xyz"), ed.NullOperator()))
    assert ed.isinvariant(hs, T, j1) and ed.isinvariant(hs, I, j1)
    assert not ed.isinvariant(hs, T, p1)


def test_translation_group_and_bit_permutation(ed):
    g = ed.translation_group((4, 3))
    assert len(g) == 12 and g[0].map == list(range(1, 13))
    hs, _ = ed.spin_half_system(12)
    bp = ed.bit_permutation(hs, ed.translation_perm((4, 3), 1, 0))
    # (x, y) -> (x+1, y): site 3 (x=3, y=0) wraps to site 0; site 4 -> 5
    assert bp[3] == 0 and bp[4] == 5
    irrep = ed.TranslationIrrep(7, 1)
    ops = irrep.symops_and_amplitudes()
    assert abs(ops[1][1] - np.exp(-2j * np.pi / 7)) < 1e-15
    # 2-bit sites move as blocks
    s3 = ed.Site([ed.State("a", 0), ed.State("b", 0), ed.State("c", 0)])
    hs3 = ed.HilbertSpace([s3, s3])
    assert ed.bit_permutation(hs3, ed.SitePermutation([2, 1])).tolist() == [2, 3, 0, 1]


# ---------------------------------------------------------------- C ABI
def _header_functions():
    with open(os.path.join(ROOT, "include", "edhip.h")) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(edhip_[a-z0-9_]+)\s*\(", txt)))


def test_abi_exports_every_declared_symbol(ed):
    from edhip import _lib

    L = _lib.lib()
    declared = _header_functions()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(L, name), name
    # the ctypes signature table covers exactly the header
    assert sorted(_lib.SIGNATURES) == declared


def test_abi_without_device_reports_device_error(ed):
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    from edhip import _lib

    h = ctypes.c_void_p()
    assert _lib.lib().edhip_create(0, ctypes.byref(h)) == _lib.EDHIP_ERR_DEVICE
    with pytest.raises(_lib.DeviceError):
        _lib.Context(0)


def test_lanczos_default_opts(ed):
    from edhip import _lib

    o = _lib.LanczosOpts()
    _lib.lib().edhip_lanczos_default_opts(ctypes.byref(o))
    assert o.nev == 1 and o.max_basis == 120 and o.tol == 1e-10 and o.seed == 20241202


def test_global_bit_flip_known(ed):
    """test_bitflip.jl:4-56: products and images of GlobalBitFlip and of its
    products with a site permutation (4 spins)."""
    b0, b1 = ed.GlobalBitFlip(), ed.GlobalBitFlip(True)
    assert b0 == ed.GlobalBitFlip(False) and b0 != ed.GlobalBitFlip(True)
    assert b0 * b0 == b0 and b0 * b1 == b1 and b1 * b0 == b1 and b1 * b1 == b0
    assert b0.isidentity() and not b1.isidentity()
    hs, _ = ed.spin_half_system(4)
    p = ed.SitePermutation([2, 3, 4, 1])
    assert ed.symmetry_apply(hs, b0, 0b0001)[0] == 0b0001
    assert ed.symmetry_apply(hs, b1, 0b0001)[0] == 0b1110
    assert ed.symmetry_apply(hs, p * b0, 0b0001)[0] == 0b0010
    assert ed.symmetry_apply(hs, p * b1, 0b0001)[0] == 0b1101
    for b in range(16):
        bp = ed.symmetry_apply(hs, p, b)[0]
        for f in (b0, b1):
            assert ed.symmetry_apply(hs, f, bp) == ed.symmetry_apply(hs, p * f, b) == ed.symmetry_apply(hs, f * p, b)
    # direct products: every (T^a, flip) once, identity first, amplitudes multiplied
    irr = ed.make_product_irrep([(b0, 1.0), (b1, -1.0)], ed.TranslationIrrep(4, 1))
    assert len(irr) == 8 and irr[0][1] == 1.0
    assert len({ed.symmetry_apply(hs, op, 0b0011)[0] for op, _ in irr}) == 4


# ------------------------------------------------ space-group characters
@pytest.mark.parametrize("L,k", [(4, (0, 0)), (4, (2, 2)), (6, (0, 0)), (6, (3, 3))])
def test_square_space_irrep_characters(ed, L, k):
    """SquareSpaceIrrep against an independent construction (conftest: C4v
    generated by the quarter turn and the axial mirror, characters from the
    generators' images): the same elements with the same characters, every
    irrep a homomorphism, and the four irreps orthonormal (class
    orthogonality: sum_g chi_a(g) conj(chi_b(g)) = |G| delta_ab)."""
    from conftest import space_group_independent

    hs, _ = ed.spin_half_system(L * L)
    names = ["A1", "A2", "B1", "B2"]
    chis = {}
    for name in names:
        ops = ed.SquareSpaceIrrep(L, k, name).symops_and_amplitudes()
        perms = np.stack([np.array(p.map) - 1 for p, _ in ops])
        chi = np.array([complex(a) for _, a in ops])
        ref_p, ref_chi = space_group_independent(L, k, name)
        ref = {tuple(p): c for p, c in zip(ref_p.tolist(), ref_chi)}
        assert len(ref) == len(ops) == 8 * L * L
        assert np.array_equal(perms[0], np.arange(L * L))
        for p, c in zip(perms.tolist(), chi):
            assert abs(ref[tuple(p)] - c) < 1e-12, name
        index = {tuple(p): i for i, p in enumerate(perms.tolist())}
        rng = np.random.default_rng(5)
        for _ in range(200):
            a, b = rng.integers(0, len(ops), 2)
            ab = tuple(perms[a][perms[b]])  # g_a after g_b
            assert abs(chi[index[ab]] - chi[a] * chi[b]) < 1e-12
        chis[name] = chi
    G = 8 * L * L
    for a in names:
        for b in names:
            s = np.vdot(chis[b], chis[a])
            assert abs(s - (G if a == b else 0)) < 1e-9


@pytest.mark.parametrize("L,k,kind", [(4, (2, 0), "C2v"), (4, (0, 2), "C2v"), (6, (3, 0), "C2v"), (4, (1, 0), "Cs"),
                                      (4, (1, 1), "Cs"), (4, (1, 2), "Cs"), (5, (2, 0), "Cs"), (5, (1, 2), "C1"),
                                      (6, (1, 2), "C1")])
def test_little_group_irreps(ed, L, k, kind):
    """Little groups of C4v at k = (L/2, 0), on mirror lines and elsewhere:
    every element fixes k (p k = k mod L) and the rest of C4v does not, the
    elements T(a, b) p form a group (closed under composition, identity
    first, each once), every irrep is a homomorphism, and the irreps are
    orthonormal: sum_g chi_a(g) conj(chi_b(g)) = |G| delta_ab."""
    hs, _ = ed.spin_half_system(L * L)
    names = {"C2v": ["A1", "A2", "B1", "B2"], "Cs": ["A'", "A''"], "C1": ["A"]}[kind]
    chis = {}
    for name in names:
        irr = ed.SquareSpaceIrrep(L, k, name)
        assert irr.kind == kind
        for m in irr.little:
            (a, b), (c, d) = m
            assert ((a * k[0] + b * k[1] - k[0]) % L, (c * k[0] + d * k[1] - k[1]) % L) == (0, 0)
        ops = irr.symops_and_amplitudes()
        assert len(ops) == irr.group_order == len(irr.little) * L * L
        perms = np.stack([np.array(p.map) - 1 for p, _ in ops])
        chi = np.array([complex(a) for _, a in ops])
        assert np.array_equal(perms[0], np.arange(L * L)) and chi[0] == 1
        index = {tuple(p): i for i, p in enumerate(perms.tolist())}
        assert len(index) == len(ops)
        for a in range(len(ops)):
            for b in np.random.default_rng(a).integers(0, len(ops), 8):
                ab = tuple(perms[a][perms[b]])  # g_a after g_b
                assert ab in index
                assert abs(chi[index[ab]] - chi[a] * chi[b]) < 1e-12
        chis[name] = chi
    G = len(chis[names[0]])
    for a in names:
        for b in names:
            assert abs(np.vdot(chis[b], chis[a]) - (G if a == b else 0)) < 1e-9
    with pytest.raises(ValueError):
        ed.SquareSpaceIrrep(L, k, "E")


@pytest.mark.parametrize("L,k", [(4, (2, 0)), (4, (1, 0)), (4, (1, 2)), (4, (1, 3))])
def test_little_group_sectors_split_the_momentum_sector(ed, oracle, L, k):
    """The symmorphic sectors of the irreps of k's little group split the
    translation sector at k (oracle reductions of the Sz = 0 sector of the
    L x L torus): their dimensions add up to its dimension."""
    n = L * L
    hs, _ = ed.spin_half_system(n)
    parent = oracle.sector_basis(n, 0)

    def dim(ops):
        perms = np.stack([np.array(p.map) - 1 for p, _ in ops])
        return oracle.symmetry_reduce(parent, n, perms, [complex(a) for _, a in ops]).basis_list.size

    total = dim(ed.TranslationIrrep((L, L), k).symops_and_amplitudes())
    kind = None
    for probe in ("A1", "A'", "A"):
        try:
            kind = ed.SquareSpaceIrrep(L, k, probe).kind
            break
        except ValueError:
            pass
    assert kind in ("C2v", "Cs", "C1")  # abelian: every irrep one-dimensional
    names = {"C2v": ["A1", "A2", "B1", "B2"], "Cs": ["A'", "A''"], "C1": ["A"]}[kind]
    parts = [dim(ed.SquareSpaceIrrep(L, k, name).symops_and_amplitudes()) for name in names]
    assert sum(parts) == total, (parts, total)
    assert all(p > 0 for p in parts)


# ------------------------------------------------ coupling invariance
def _fake_reduced(ed, hs, ops):
    """A ReducedHilbertSpaceRepresentation shell carrying only its group (the
    check below reads nothing else); no device call."""
    r = object.__new__(ed.ReducedHilbertSpaceRepresentation)
    r.hilbert_space = hs
    r.symops_and_amplitudes = list(ops)
    r._handle = None
    return r


def test_linear_couplings_invariance_check(ed):
    """On a symmetry-reduced space a coupling operator must commute with every
    group element: per-bond operators are rejected, their symmetric sums pass
    (the reduced element formula applies an operator to the representative
    of each row only)."""
    from edhip.grad import _check_invariant

    L = 8
    hs, s = ed.spin_half_system(L)
    bonds = [(i, (i + 1) % L) for i in range(L)]
    per_bond = [ed.heisenberg_bond(hs, i, j) for i, j in bonds]
    total = ed.simplify(sum(per_bond, ed.NullOperator()))
    zz = ed.simplify(sum((s(i, "z") * s(j, "z") for i, j in bonds), ed.NullOperator()))
    field = ed.simplify(sum((s(i, "z") for i in range(L)), ed.NullOperator()))
    trans = ed.TranslationIrrep(L, 1).symops_and_amplitudes()
    space = _fake_reduced(ed, hs, trans)
    _check_invariant(space, [total, zz, field])
    with pytest.raises(ed.ArgumentError):
        _check_invariant(space, per_bond)
    flip = ed.make_product_irrep([(ed.GlobalBitFlip(False), 1.0), (ed.GlobalBitFlip(True), 1.0)],
                                 ed.TranslationIrrep(L, 0))
    fspace = _fake_reduced(ed, hs, flip)
    _check_invariant(fspace, [total, zz])
    with pytest.raises(ed.ArgumentError):
        _check_invariant(fspace, [field])  # sum_i S^z_i changes sign under the flip
    # an unreduced space takes any operator (no group)
    _check_invariant(object(), per_bond)
