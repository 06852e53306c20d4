"""Device enumeration of sectors of general sites (several states per site,
quantum-number tuples) vs the reference's fixtures and the oracle's
hs_get_basis_list (hilbert_space_representation.jl:109-206), then lookups,
products and eigenvalues on such sectors."""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as sla

pytestmark = pytest.mark.gpu


def _space(ed, site_qns):
    sites = []
    for states in site_qns:
        sites.append(ed.Site([ed.State(f"s{k}", tuple(q)) for k, q in enumerate(states)]))
    return ed.HilbertSpace(sites)


def _spin_one_qns(n):
    return [[(1,), (0,), (-1,)] for _ in range(n)]


def _tj_qns(n):
    return [[(0, 0), (1, 1), (1, -1)] for _ in range(n)]


def test_tj_fixture(ed, golden):
    """test_hilbert_space_representation.jl:73-128: one three-state site and two
    two-state sites, (N, 2Sz) tuples; every sector, the full space, two sectors
    together and an empty allowed set."""
    ka = golden["known_answers"]["tj_sectors"]
    site_qns = [[tuple(s[1]) for s in site] for site in ka["sites"]]
    hs = _space(ed, site_qns)
    assert ed.represent(hs).basis_list.tolist() == ka["full_basis"]
    for key, basis in ka["sector_basis"].items():
        q = tuple(int(x) for x in key.split(","))
        assert ed.represent(ed.HilbertSpaceSector(hs, q)).basis_list.tolist() == basis
    two = ka["two_sector_basis"]
    hss = ed.HilbertSpaceSector(hs, [tuple(q) for q in two["sectors"]])
    assert ed.represent(hss).basis_list.tolist() == two["basis"]
    empty = ed.represent(ed.HilbertSpaceSector(hs, [(7, 7)]))
    assert empty.dimension == 0 and empty.basis_list.tolist() == []


@pytest.mark.parametrize("n", [1, 5, 10])
def test_spin_one_sectors_vs_oracle(ed, oracle, n):
    qns = _spin_one_qns(n)
    hs = _space(ed, qns)
    total = 0
    for q in oracle.quantum_number_sectors(qns):
        b = ed.represent(ed.HilbertSpaceSector(hs, q)).basis_list
        assert np.array_equal(b, oracle.hs_get_basis_list(qns, {q}))
        total += b.size
    assert total == 3 ** n
    full = ed.represent(hs).basis_list
    assert np.array_equal(full, oracle.hs_get_basis_list(qns, set(oracle.quantum_number_sectors(qns))))


def test_tj_chain_sectors_vs_oracle(ed, oracle):
    n = 10
    qns = _tj_qns(n)
    hs = _space(ed, qns)
    for q in [(0, 0), (10, 0), (9, 1), (6, -2), (5, 5)]:
        hsr = ed.represent(ed.HilbertSpaceSector(hs, q))
        assert np.array_equal(hsr.basis_list, oracle.hs_get_basis_list(qns, {q})), q
    both = ed.represent(ed.HilbertSpaceSector(hs, [(8, 0), (8, 2), (7, -1)])).basis_list
    assert np.array_equal(both, oracle.hs_get_basis_list(qns, {(8, 0), (8, 2), (7, -1)}))


def test_mixed_sites_vs_oracle(ed, oracle):
    """Sites of 1, 2, 3, 5 and 4 states (0, 1, 2, 3 and 2 bits), two-component
    quantum numbers drawn from a seed, with repeats between states."""
    rng = np.random.default_rng(11)
    sizes = [3, 2, 5, 1, 4, 3, 2, 5, 3, 4]
    qns = [[tuple(int(x) for x in rng.integers(-2, 3, size=2)) for _ in range(k)] for k in sizes]
    hs = _space(ed, qns)
    sectors = oracle.quantum_number_sectors(qns)
    for q in sectors[:: max(1, len(sectors) // 12)]:
        assert np.array_equal(ed.represent(ed.HilbertSpaceSector(hs, q)).basis_list, oracle.hs_get_basis_list(qns, {q}))
    pick = set(sectors[1::7])
    assert np.array_equal(ed.represent(ed.HilbertSpaceSector(hs, list(pick))).basis_list,
                          oracle.hs_get_basis_list(qns, pick))
    full = ed.represent(hs).basis_list
    assert full.size == math.prod(sizes)
    assert np.array_equal(full, oracle.hs_get_basis_list(qns, set(sectors)))


def test_spin_one_large_sector_and_lookup(ed, oracle):
    """16 spin-1 sites, Sz = 0 (5,196,627 states = the central trinomial
    coefficient): against the oracle, and every word's lookup is its index."""
    n = 16
    qns = _spin_one_qns(n)
    hsr = ed.represent(ed.HilbertSpaceSector(_space(ed, qns), 0))
    assert hsr.dimension == sum(math.comb(n, k) * math.comb(n - k, k) for k in range(n // 2 + 1))
    b = hsr.basis_list
    assert np.array_equal(b, oracle.hs_get_basis_list(qns, {(0,)}))
    sample = b[::37]
    assert np.array_equal(hsr.lookup(sample), np.arange(0, b.size, 37))
    # the state index 3 never appears in a 2-bit field of a three-state site
    assert np.all(hsr.lookup(np.array([3, 3 << 30], np.uint64)) == -1)


def _spin_one_ops(ed, hs, n):
    """S_i . S_j for spin-1 sites with states (+1, 0, -1) = indices (0, 1, 2):
    Sz Sz + (S+ S- + S- S+)/2, S+|m> = sqrt(2)|m+1> for m = 0, -1."""
    def sz(i):
        return ed.pure_operator(hs, i, 0, 0, 1.0) + ed.pure_operator(hs, i, 2, 2, -1.0)

    def sp(i):
        r2 = math.sqrt(2.0)
        return ed.pure_operator(hs, i, 0, 1, r2) + ed.pure_operator(hs, i, 1, 2, r2)

    def sm(i):
        r2 = math.sqrt(2.0)
        return ed.pure_operator(hs, i, 1, 0, r2) + ed.pure_operator(hs, i, 2, 1, r2)

    terms = ed.NullOperator()
    for i in range(n):
        j = (i + 1) % n
        terms = terms + sz(i) * sz(j) + 0.5 * (sp(i) * sm(j) + sm(i) * sp(j))
    return ed.simplify(terms)


def _dense_spin_one_chain(n):
    """The same Hamiltonian from Kronecker products, basis index = sum s_i 3^i
    with s = 0, 1, 2 for m = +1, 0, -1."""
    sz = np.diag([1.0, 0.0, -1.0])
    sp = np.zeros((3, 3))
    sp[0, 1] = sp[1, 2] = math.sqrt(2.0)
    sm = sp.T

    def site(op, i):
        m = np.ones((1, 1))
        for k in reversed(range(n)):  # site n-1 is the most significant
            m = np.kron(m, op if k == i else np.eye(3))
        return m

    H = np.zeros((3 ** n, 3 ** n))
    for i in range(n):
        j = (i + 1) % n
        H += site(sz, i) @ site(sz, j) + 0.5 * (site(sp, i) @ site(sm, j) + site(sm, i) @ site(sp, j))
    return H


def test_spin_one_chain_spectrum(ed):
    """Spin-1 Heisenberg ring of 6 sites: the union of the Sz-sector spectra of
    the device operator equals the spectrum of the Kronecker-product matrix,
    and the lowest level sits in Sz = 0."""
    n = 6
    qns = _spin_one_qns(n)
    hs = _space(ed, qns)
    H = _spin_one_ops(ed, hs, n)
    ref = np.linalg.eigvalsh(_dense_spin_one_chain(n))
    got = []
    for q in range(-n, n + 1):
        opr = ed.represent(ed.represent(ed.HilbertSpaceSector(hs, q)), H)
        M = opr.sparse().toarray()
        assert np.abs(M - M.conj().T).max() < 1e-13
        got.append(np.linalg.eigvalsh(M))
    allv = np.sort(np.concatenate(got))
    assert allv.size == 3 ** n
    assert np.abs(allv - ref).max() < 1e-11
    assert abs(got[n].min() - ref.min()) < 1e-11


def test_spin_one_matvec_and_lanczos(ed, oracle):
    """12-site spin-1 ring, Sz = 0 (73,789 states): products against the oracle
    (FP64, 1e-13 relative) and the device Lanczos E0 against ARPACK on the
    oracle's sparse matrix (1e-12 relative)."""
    n = 12
    qns = _spin_one_qns(n)
    hs = _space(ed, qns)
    H = _spin_one_ops(ed, hs, n)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    words = oracle.hs_get_basis_list(qns, {(0,)})
    assert np.array_equal(hsr.basis_list, words)
    opr = ed.represent(hsr, H)
    o = oracle.Opr(words, ed.terms_arrays(H))
    rng = np.random.default_rng(5)
    x = rng.standard_normal(words.size)
    y = opr @ x
    yo = o.matvec(x).real
    assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()
    res = ed.eigs(opr, 1, tol=1e-12)
    e_ref = sla.eigsh(o.scipy_sparse(), k=1, which="SA", tol=1e-13)[0][0]
    assert abs(res.values[0] - e_ref) <= 1e-12 * abs(e_ref)


def test_tj_reduce_translations(ed, oracle):
    """Translations of a 10-site t-J ring act on 2-bit site fields: the reduced
    basis of a QN sector built on the device matches the oracle's reduction of
    the oracle's sector list."""
    n = 10
    qns = _tj_qns(n)
    hs = _space(ed, qns)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, (8, 0)))
    words = oracle.hs_get_basis_list(qns, {(8, 0)})
    assert np.array_equal(hsr.basis_list, words)
    ops = [(ed.translation_perm(n, a), np.exp(-2j * np.pi * 3 * a / n)) for a in range(n)]
    r = ed.symmetry_reduce(hsr, ops)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in ops])
    red = oracle.symmetry_reduce(words, hs.bitwidth(), perms, [a for _, a in ops])
    assert np.array_equal(r.basis_list, red.basis_list)
    assert np.array_equal(r.basis_mapping_index, red.basis_mapping_index)
    assert np.abs(r.basis_mapping_amplitude - red.basis_mapping_amplitude).max() < 1e-14


def test_qn_argument_errors(ed):
    s3 = ed.Site([ed.State("a", 0), ed.State("b", 1), ed.State("c", 2)])
    hs = ed.HilbertSpace([s3] * 32)  # 64 bits: UInt64 is not enough
    with pytest.raises(ed.ArgumentError):
        ed.represent(ed.HilbertSpaceSector(hs, 3))
    big = ed.HilbertSpace([s3] * 31)  # 62 bits, more than 2^32 states in one sector
    with pytest.raises(ed.ArgumentError):
        ed.represent(ed.HilbertSpaceSector(big, 31))
