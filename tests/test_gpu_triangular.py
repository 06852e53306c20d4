"""Triangular-lattice clusters with their space groups (translations x the
little group of 6mm), the setting of the reference's
examples/spinhalf_triangular.jl: groups of 144 and 324 elements (288 with
the spin flip), whose 4-bit permutation tables are too large for LDS and are
read from global memory.
Reduction, products and spectra against the oracle."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# lattice coordinates: a1 = (1, 0), a2 = (-1/2, sqrt(3)/2); columns are the images of a1, a2
R60 = np.array([[1, -1], [1, 0]])
MIRROR = np.array([[-1, 1], [0, 1]])


def _cluster(shape):
    """Sites of Z^2 / S Z^2 (S columns: the supercell vectors) and the
    space-group elements r -> P r + t (mod S) that map the cluster to itself."""
    S = np.array(shape, dtype=np.int64)
    n = abs(round(np.linalg.det(S)))
    Sinv = np.linalg.inv(S)

    def wrap(r):
        f = np.floor(Sinv @ r + 1e-9)
        return tuple(int(v) for v in (r - S @ f).round().astype(np.int64))

    sites = sorted({wrap(np.array([i, j])) for i in range(-2 * n, 2 * n) for j in range(-2 * n, 2 * n)})
    assert len(sites) == n
    index = {r: k for k, r in enumerate(sites)}
    points = []
    P = np.eye(2, dtype=np.int64)
    for _ in range(6):
        for Q in (P, P @ MIRROR):
            U = Sinv @ Q @ S
            if np.allclose(U, U.round()):
                points.append(Q.copy())
        P = R60 @ P
    elements = []  # the identity first: P = 1 first, the zero translation first
    shifts = [(0, 0)] + [t for t in sites if t != (0, 0)]
    for Q in points:
        for t in shifts:
            m = [index[wrap(Q @ np.array(r) + np.array(t))] + 1 for r in sites]
            elements.append((m, int(round(np.linalg.det(Q)))))
    bonds = set()
    for r in sites:
        for d in ((1, 0), (1, 1), (0, 1)):
            a, b = index[r], index[wrap(np.array(r) + np.array(d))]
            bonds.add((min(a, b), max(a, b)))
    return n, elements, len(points), sorted(bonds)


def _ops(ed, elements, irrep, flip):
    """(operation, character) for the A1 (irrep 0) or A2 (irrep 1, the sign of
    det P) component at k = 0, optionally times the spin flip with sign flip."""
    base = [(ed.SitePermutation(m), 1.0 if irrep == 0 else float(det)) for m, det in elements]
    if flip is None:
        return base
    return ed.make_product_irrep([(ed.GlobalBitFlip(False), 1.0), (ed.GlobalBitFlip(True), float(flip))], base)


def _heis(ed, hs, bonds):
    return ed.simplify(sum((ed.heisenberg_bond(hs, i, j, 1.0) for i, j in bonds), ed.NullOperator()))


@pytest.mark.parametrize("irrep,flip", [(0, None), (1, None), (0, 1), (0, -1), (1, 1)])
def test_triangular12_vs_oracle(ed, oracle, irrep, flip):
    """12-site cluster (S = [[2, 2], [-2, 4]], full 6mm little group: 144
    elements, 288 with the flip): basis, mapping and amplitudes against the
    oracle's symmetry_reduce, products against the oracle, and the sector's
    spectrum against the dense matrix of the oracle's operator."""
    n, elements, n_points, bonds = _cluster([[2, 2], [-2, 4]])
    assert n == 12 and n_points == 12 and len(bonds) == 36
    hs, _ = ed.spin_half_system(n)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    ops = _ops(ed, elements, irrep, flip)
    r = ed.symmetry_reduce(hsr, ops)
    assert r.canonicaliser() == "generic"
    perms, chars, flips = [], [], []
    for op, ch in ops:
        p, fl = ed.representation._perm_and_flip(hs, op)
        perms.append(p)
        flips.append(fl)
        chars.append(ch)
    red = oracle.symmetry_reduce(oracle.sector_basis(n, 0), n, np.stack(perms), chars,
                                 flips=np.array(flips, np.uint64) if flip is not None else None)
    assert np.array_equal(r.basis_list, red.basis_list)
    assert np.array_equal(r.basis_mapping_index, red.basis_mapping_index)
    assert np.abs(r.basis_mapping_amplitude - red.basis_mapping_amplitude).max() < 1e-13
    H = _heis(ed, hs, bonds)
    opr = ed.represent(r, H)
    o = oracle.Opr(red, ed.terms_arrays(H))
    x = np.random.default_rng(irrep + 3).standard_normal(opr.dimension)
    assert np.abs(opr @ x - o.matvec(x).real).max() < 1e-12 * max(1.0, np.abs(x).max())
    M = opr.Matrix()
    assert np.abs(M - o.dense().real).max() < 1e-13
    assert np.abs(M - M.T).max() < 1e-13


def test_triangular12_dimensions_by_characters(ed):
    """The dimension of each 1-D sector of the 288-element group (A1, A2 at
    k = 0, times the even or odd spin flip) is the multiplicity of its
    character in the permutation representation on the Sz = 0 words:
    (1/|G|) sum_g chi(g) #(words fixed by g)."""
    n, elements, _, _ = _cluster([[2, 2], [-2, 4]])
    hs, _ = ed.spin_half_system(n)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    words = hsr.basis_list

    def fixed(m, flipped):
        perm = np.array(m) - 1
        cnt = 0
        for w in words.tolist():
            img = 0
            for i in range(n):
                if (w >> i) & 1:
                    img |= 1 << int(perm[i])
            if flipped:
                img ^= (1 << n) - 1
            cnt += img == w
        return cnt

    fix = [(fixed(m, False), fixed(m, True), det) for m, det in elements]
    G = 2 * len(elements)
    for irrep in (0, 1):
        for sign in (1, -1):
            chi = lambda det: 1.0 if irrep == 0 else float(det)
            expect = sum(chi(det) * (a + sign * b) for a, b, det in fix) / G
            r = ed.symmetry_reduce(hsr, _ops(ed, elements, irrep, sign))
            assert r.dimension == round(expect)


def test_triangular27_orbits(ed):
    """27-site cluster of examples/spinhalf_triangular.jl (S = [[3, 3], [-3, 6]],
    324 elements): the A1 sector of Sz = 1/2; representatives ascending and
    minimal in their orbits, orbit sizes divide 324 and add up to C(27, 13)."""
    n, elements, n_points, _ = _cluster([[3, 3], [-3, 6]])
    assert n == 27 and n_points == 12
    hs, _ = ed.spin_half_system(n)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 1))
    r = ed.symmetry_reduce(hsr, _ops(ed, elements, 0, None))
    orb = r.orbit_sizes.astype(np.int64)
    assert int(orb.sum()) == math.comb(27, 13)
    assert np.all(324 % orb == 0)
    b = r.basis_list
    assert np.all(b[1:] > b[:-1])
    perms = [np.array(m) - 1 for m, _ in elements]
    for w in b[:: max(1, b.size // 200)].tolist():
        for p in perms[:: 7]:
            img = 0
            for i in range(n):
                if (w >> i) & 1:
                    img |= 1 << int(p[i])
            assert img >= w
