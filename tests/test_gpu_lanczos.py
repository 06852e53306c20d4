"""Lanczos eigensolver on the device (GPU).

FP64 mode: eigenvalues within 1e-12 relative of the ARPACK (scipy eigsh)
spectrum of the oracle's sparse matrix.  Mixed mode (FP32 Krylov storage +
FP64 refinement): within 1e-10 relative of the FP64 path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _heis(ed, hs, bonds, w=None):
    return ed.simplify(sum((ed.heisenberg_bond(hs, i, j, 1.0 if w is None else w[k]) for k, (i, j) in enumerate(bonds)), ed.NullOperator()))


def test_chain16_unreduced_known(ed, golden):
    L = 16
    hs, _ = ed.spin_half_system(L)
    H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)])
    opr = ed.represent(ed.represent(ed.HilbertSpaceSector(hs, 0)), H)
    ref = np.array(golden["eigenvalues_brute_force"]["chain16"])
    r = ed.eigs(opr, 4, tol=1e-12, max_basis=200)
    assert r.converged
    assert np.all(np.abs(r.values - ref) < 1e-11 * np.abs(ref))
    assert abs(r.values[0] - (-7.14229636)) < 1e-8
    rm = ed.eigs(opr, 4, tol=1e-11, mode="mixed", max_basis=200)
    assert np.all(np.abs(rm.values - ref) < 1e-10 * np.abs(ref))
    assert rm.matvecs_fp64 > 0


def test_thick_restart(ed, golden):
    """A basis of 24 vectors forces thick restarts; same spectrum."""
    L = 16
    hs, _ = ed.spin_half_system(L)
    H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)])
    opr = ed.represent(ed.represent(ed.HilbertSpaceSector(hs, 0)), H)
    ref = np.array(golden["eigenvalues_brute_force"]["chain16"])[:2]
    r = ed.eigs(opr, 2, tol=1e-11, max_basis=24, max_restarts=400)
    assert r.restarts > 0 and r.converged
    assert np.all(np.abs(r.values - ref) < 1e-10 * np.abs(ref))


def test_square4x4_over_k(ed, golden):
    hs, _ = ed.spin_half_system(16)
    j1, _ = ed.square_bonds(4, 4)
    H = _heis(ed, hs, j1)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    e = min(ed.eigs(ed.represent(ed.symmetry_reduce(hsr, ed.TranslationIrrep((4, 4), (kx, ky))), H), 1, tol=1e-12).values[0]
            for kx in range(4) for ky in range(4))
    ref = golden["eigenvalues_brute_force"]["square4x4"][0]
    assert abs(e - ref) < 1e-11 * abs(ref)


def test_eigenvectors(ed):
    L = 14
    hs, _ = ed.spin_half_system(L)
    H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)])
    opr = ed.represent(ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ed.TranslationIrrep(L, 3)), H)
    r = ed.eigs(opr, 2, tol=1e-12, return_vectors=True)
    V = r.vectors.cpu().numpy()
    for i in range(2):
        v = V[i]
        assert abs(np.linalg.norm(v) - 1) < 1e-10
        assert np.linalg.norm(opr @ v - r.values[i] * v) < 1e-9
    assert abs(np.vdot(V[0], V[1])) < 1e-9


@pytest.mark.parametrize("name,shape,k,nev", [("chain24_k3", (24, 1), (3, 0), 4), ("torus4x6_pi0", (4, 6), (2, 0), 1)])
@pytest.mark.parametrize("reorth", ["auto", "partial"])
def test_full_size_configs_vs_arpack(ed, oracle, name, shape, k, nev, reorth):
    """The 24-site configurations: FP64 Lanczos vs ARPACK on the oracle's
    sparse matrix; mixed precision vs FP64 at 1e-10 relative; with full
    (auto at this size) and partial reorthogonalisation."""
    n = shape[0] * shape[1]
    hs, _ = ed.spin_half_system(n)
    if shape[1] == 1:
        H = _heis(ed, hs, [(i, (i + 1) % n) for i in range(n)])
    else:
        j1, j2 = ed.square_bonds(*shape)
        H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    irrep = ed.TranslationIrrep(shape, k)
    opr = ed.represent(ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), irrep), H)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in irrep.symops_and_amplitudes()])
    red = oracle.symmetry_reduce(oracle.sector_basis(n, 0), n, perms, [a for _, a in irrep.symops_and_amplitudes()])
    ref = oracle.lowest_eigenvalues(oracle.Opr(red, ed.terms_arrays(H)), nev)
    r64 = ed.eigs(opr, nev, tol=1e-12, max_basis=200, reorth=reorth)
    assert np.all(np.abs(r64.values - ref) < 1e-12 * np.abs(ref))
    rmx = ed.eigs(opr, nev, tol=1e-11, mode="mixed", max_basis=200, criterion="eigenvalue", reorth=reorth)
    assert np.all(np.abs(rmx.values - r64.values) < 1e-10 * np.abs(r64.values))
    if reorth == "partial":
        assert r64.full_reorth_steps < r64.iterations // 2
    else:
        assert r64.full_reorth_steps == r64.iterations


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_reorth_adaptive_matches_twice(ed, golden, mode):
    """One Gram-Schmidt pass per step (second pass only on heavy cancellation)
    gives the same four lowest levels as always-twice CGS2, within the
    precision bar of the mode, with thick restarts in play; the eigenvectors
    stay orthonormal."""
    L = 16
    hs, _ = ed.spin_half_system(L)
    H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)])
    opr = ed.represent(ed.represent(ed.HilbertSpaceSector(hs, 0)), H)
    ref = np.array(golden["eigenvalues_brute_force"]["chain16"])
    kw = dict(tol=1e-12 if mode == "fp64" else 1e-11, mode=mode, max_basis=40, max_restarts=400)
    a = ed.eigs(opr, 4, reorth="adaptive", return_vectors=True, **kw)
    t = ed.eigs(opr, 4, reorth="twice", **kw)
    p = ed.eigs(opr, 4, reorth="partial", return_vectors=True, **kw)
    bar = 1e-11 if mode == "fp64" else 1e-10
    assert a.converged and t.converged and p.converged and a.restarts > 0
    assert np.all(np.abs(a.values - ref) < bar * np.abs(ref))
    assert np.all(np.abs(t.values - ref) < bar * np.abs(ref))
    assert np.all(np.abs(p.values - ref) < bar * np.abs(ref))
    assert a.second_passes < t.second_passes
    V = a.vectors.cpu().numpy()
    assert np.abs(V @ V.T - np.eye(4)).max() < 1e-9
    V = p.vectors.cpu().numpy()
    assert np.abs(V @ V.T - np.eye(4)).max() < 1e-9
    with pytest.raises(ed.ArgumentError):
        ed.eigs(opr, 1, reorth="never")


def test_workspace_reuse_repeatable(ed):
    """Calls that reuse the context's workspace (growing it, shrinking the
    request, after a release) give bitwise the same results."""
    L = 14
    hs, _ = ed.spin_half_system(L)
    H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)])
    opr = ed.represent(ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ed.TranslationIrrep(L, 2)), H)
    runs = []
    for mode, basis in (("mixed", 30), ("fp64", 60), ("mixed", 30), ("fp64", 60)):
        runs.append(ed.eigs(opr, 2, tol=1e-11, mode=mode, max_basis=basis, max_restarts=200).values)
    ed.release_workspace(opr.ctx)
    runs.append(ed.eigs(opr, 2, tol=1e-11, mode="fp64", max_basis=60, max_restarts=200).values)
    assert np.array_equal(runs[0], runs[2])
    assert np.array_equal(runs[1], runs[3]) and np.array_equal(runs[1], runs[4])
    assert np.all(np.abs(runs[0] - runs[1]) < 1e-10 * np.abs(runs[1]))


def test_restart_in_row_chunks_space_group_6x6(ed):
    """6x6 J1-J2, translations x C4v, A1 (31,566,122 states): a 12-vector
    basis forces thick restarts whose combination runs in two row chunks of
    the restart scratch; E0 = -18.13714754 as in the translation sector."""
    hs, _ = ed.spin_half_system(36)
    j1, j2 = ed.square_bonds(6, 6)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    r = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ed.SquareSpaceIrrep(6, (0, 0), "A1"))
    assert r.dimension == 31566122
    opr = ed.represent(r, H)
    res = ed.eigs(opr, 1, tol=1e-10, max_basis=12, max_restarts=400, criterion="eigenvalue")
    assert res.converged and res.restarts > 0
    assert abs(res.values[0] - (-18.13714754)) < 1e-8


def _zz_chain(ed, L):
    hs, s = ed.spin_half_system(L)
    zz = ed.simplify(sum((s(i, "z") * s((i + 1) % L, "z") for i in range(L)), ed.NullOperator()))
    return hs, zz


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("max_basis", [12, 120])
@pytest.mark.parametrize("reorth", ["auto", "partial"])
def test_invariant_subspace_degenerate_levels(ed, mode, max_basis, reorth):
    """A diagonal operator: the Krylov space of one start vector holds one copy
    of each distinct eigenvalue (4 here), so the lowest three levels with
    their multiplicities (-8, -8, -4 for sum_i sigma^z_i sigma^z_{i+1} on the
    8-site ring, Sz = 0: the two Neel states first) need random directions
    after each invariant subspace; the eigenvalues are the sorted diagonal.
    max_basis 12 also restarts; 120 exceeds the dimension (70), where the
    basis spans the whole space."""
    L = 8
    hs, zz = _zz_chain(ed, L)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    opr = ed.represent(hsr, zz)
    words = hsr.basis_list
    diag = np.zeros(len(words))
    for i in range(L):
        bi = ((words >> np.uint64(i)) & np.uint64(1)).astype(np.int64)
        bj = ((words >> np.uint64((i + 1) % L)) & np.uint64(1)).astype(np.int64)
        diag += 1 - 2 * (bi ^ bj)
    ref = np.sort(diag)[:3]
    assert ref.tolist() == [-8.0, -8.0, -4.0]
    r = ed.eigs(opr, 3, tol=1e-12, mode=mode, max_basis=max_basis, max_restarts=400, return_vectors=True,
                reorth=reorth)
    assert r.converged
    assert np.all(np.abs(r.values - ref) < 1e-11), r.values
    V = r.vectors.cpu().numpy()
    assert np.abs(V @ V.T - np.eye(3)).max() < 1e-9
    for i in range(3):
        assert np.linalg.norm(diag * V[i] - r.values[i] * V[i]) < 1e-8


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("reorth", ["auto", "partial"])
def test_invariant_subspace_zero_operator(ed, mode, reorth):
    """sum_i sigma^z_i vanishes on the Sz = 0 sector: every Krylov block has one
    vector; four zero eigenvalues come back converged."""
    L = 10
    hs, s = ed.spin_half_system(L)
    op = sum((s(i, "z") for i in range(L)), ed.NullOperator())
    opr = ed.represent(ed.represent(ed.HilbertSpaceSector(hs, 0)), op)
    r = ed.eigs(opr, 4, tol=1e-12, mode=mode, max_basis=20, max_restarts=50, reorth=reorth)
    assert r.converged and np.all(np.abs(r.values) < 1e-12)
