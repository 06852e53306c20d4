"""Vector maps between a parent sector and its symmetry reduction on the
device, symmetry_reduce(rhsr, v) / symmetry_unreduce(rhsr, v)
(symmetry_reduce.jl:29-84,208-225), vs the oracle's restatement over the
basis_mapping arrays (GPU).  FP64/complex128 within 1e-12."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reduced(ed, oracle, hs, hsr, parent, ops, n):
    r = ed.symmetry_reduce(hsr, ops)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in ops])
    red = oracle.symmetry_reduce(parent, n, perms, [a for _, a in ops])
    assert np.array_equal(r.basis_list, red.basis_list)
    return r, red


def _check_maps(ed, oracle, r, red, seed, cx=True):
    rng = np.random.default_rng(seed)
    npar, nred = red.parent.size, red.dimension
    large = rng.standard_normal(npar) + (1j * rng.standard_normal(npar) if cx else 0)
    small = rng.standard_normal(nred) + (1j * rng.standard_normal(nred) if cx else 0)
    got_s = ed.symmetry_reduce(r, large)
    got_l = ed.symmetry_unreduce(r, small)
    ref_s = oracle.vector_reduce(red, large)
    ref_l = oracle.vector_unreduce(red, small)
    assert got_s.shape == (nred,) and got_l.shape == (npar,)
    assert np.abs(got_s - ref_s).max(initial=0) < 1e-12 * max(1.0, np.abs(ref_s).max(initial=0))
    assert np.abs(got_l - ref_l).max(initial=0) < 1e-12
    # reduce(unreduce(v)) == v (U^dagger U = 1 on the reduced space)
    back = ed.symmetry_reduce(r, got_l)
    assert np.abs(back - small).max(initial=0) < 1e-12
    return got_s, got_l


@pytest.mark.parametrize("k", [0, 1, 3, 6])
def test_chain12_translation(ed, oracle, k):
    hs, _ = ed.spin_half_system(12)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    r, red = _reduced(ed, oracle, hs, hsr, oracle.sector_basis(12, 0), ed.TranslationIrrep(12, k).symops_and_amplitudes(), 12)
    assert r.canonicaliser() == "torus"
    got_s, got_l = _check_maps(ed, oracle, r, red, 10 + k)
    assert np.iscomplexobj(got_s) and np.iscomplexobj(got_l)


def test_square4x4_torus(ed, oracle):
    hs, _ = ed.spin_half_system(16)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    parent = oracle.sector_basis(16, 0)
    for k in [(0, 0), (1, 2), (2, 0), (3, 3)]:
        r, red = _reduced(ed, oracle, hs, hsr, parent, ed.TranslationIrrep((4, 4), k).symops_and_amplitudes(), 16)
        _check_maps(ed, oracle, r, red, 7)


@pytest.mark.parametrize("step,canon", [(1, "torus"), (2, "generic")])
@pytest.mark.parametrize("parity", [1, -1])
def test_dihedral_real(ed, oracle, parity, step, canon):
    """Real characters given as floats: real vectors in, real vectors out
    (promote_type(Float64, Float64)).  All translations x inversion run on the
    torus canonicaliser (two cosets of the translations); even translations x
    inversion on the generic one."""
    L = 12
    hs, _ = ed.spin_half_system(L)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    T = ed.translation_group(L)
    inv = ed.SitePermutation([((-i) % L) + 1 for i in range(L)])
    ops = [(T[a], 1.0) for a in range(0, L, step)] + [(T[a] * inv, float(parity)) for a in range(0, L, step)]
    r, red = _reduced(ed, oracle, hs, hsr, oracle.sector_basis(L, 0), ops, L)
    assert r.canonicaliser() == canon
    got_s, got_l = _check_maps(ed, oracle, r, red, 3, cx=False)
    assert got_s.dtype == np.float64 and got_l.dtype == np.float64


def test_chain24_k3_full_size(ed, oracle):
    """The 24-site ring at k = 3 (2,704,156 parent words, 112,640 representatives)."""
    n = 24
    hs, _ = ed.spin_half_system(n)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    r, red = _reduced(ed, oracle, hs, hsr, oracle.sector_basis(n, 0), ed.TranslationIrrep(n, 3).symops_and_amplitudes(), n)
    _check_maps(ed, oracle, r, red, 24)


def test_eigenvector_unreduces_to_sector_eigenvector(ed):
    """An eigenvector of the reduced Hamiltonian, unreduced, is an eigenvector
    of the unreduced sector Hamiltonian with the same eigenvalue."""
    import torch

    L = 14
    hs, _ = ed.spin_half_system(L)
    H = ed.simplify(sum((ed.heisenberg_bond(hs, i, (i + 1) % L, 1.0) for i in range(L)), ed.NullOperator()))
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    r = ed.symmetry_reduce(hsr, ed.TranslationIrrep(L, 2))
    res = ed.eigs(ed.represent(r, H), 1, tol=1e-12, return_vectors=True)
    v = res.vectors[0]
    big = ed.symmetry_unreduce(r, v)
    assert torch.is_tensor(big) and big.is_cuda and big.dtype == torch.complex128
    Hp = ed.represent(hsr, H)
    big_h = big.cpu().numpy()
    assert abs(np.linalg.norm(big_h) - 1) < 1e-10
    assert np.linalg.norm(Hp @ big_h - res.values[0] * big_h) < 1e-9
    back = ed.symmetry_reduce(r, big)
    assert torch.is_tensor(back) and torch.allclose(back, v.to(torch.complex128), atol=1e-12)


def test_dimension_mismatch(ed):
    hs, _ = ed.spin_half_system(8)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    r = ed.symmetry_reduce(hsr, ed.TranslationIrrep(8, 1))
    with pytest.raises(ed.DimensionMismatch):
        ed.symmetry_reduce(r, np.zeros(hsr.dimension + 1))
    with pytest.raises(ed.DimensionMismatch):
        ed.symmetry_unreduce(r, np.zeros(r.dimension + 1))
