"""The 6x6 J1-J2 k = (0,0) translation sector (252,091,362 representatives),
the benchmarked workload, through the products the benchmark times: the
translation-only torus kernels with the 6x6 shape fixed at compile time,
FP32 products reading x spread over word positions and FP64 products with
the blocked rank index, row form and column form (GPU).

No oracle fits this sector (its parent arrays hold 9.08e9 words), so the
checks are independent of the product's own inputs and of its kernels:
  * the group and the terms come from conftest's helpers (site maps of
    T(a, b), hand-written PureOperator terms), not from the package;
  * sampled rows of y = H x are compared with host sums built from the
    reduced element formula of reduced_operator_representation.jl:57-85
    (element a * U[b'] / U[b] with U = 1 / sqrt(orbit size) at k = 0, the
    image's representative the minimum over the 36 translated words);
  * the translation-sector ground-state energy (FP64 and mixed) equals the
    space-group A1 energy, computed by another kernel instance (the coset
    canonicaliser) on another basis.
"""
import numpy as np
import pytest

from conftest import (heisenberg_terms_pauli, permute_words, space_group_independent, square_bonds_independent,
                      translation_group_perms)

pytestmark = pytest.mark.gpu

L = 6
D6 = 252091362
E0_REF = -18.13714754  # 6x6 J1-J2 (J2 = 0.5) ground state, S.S units


def _terms(ed):
    j1, j2 = square_bonds_independent(L, L)
    mask, row, col, amp = heisenberg_terms_pauli(j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    amp = 0.25 * amp.real  # S.S units
    op = ed.SumOperator([ed.PureOperator(int(m), int(r), int(c), float(a)) for m, r, c, a in zip(mask, row, col, amp)])
    return op, (mask, row, col, amp)


@pytest.fixture(scope="module")
def sector(ed):
    hs, _ = ed.spin_half_system(L * L)
    T = translation_group_perms(L, L)
    ops = [(ed.SitePermutation([int(j) + 1 for j in p]), 1.0) for p in T]
    rh = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ops)
    op, terms = _terms(ed)
    opr = ed.represent(rh, op)
    yield hs, T, rh, opr, terms
    del opr, rh
    ed.release_workspace()


def _canonical(words, T):
    """Representative (minimum over the translated words) and orbit size."""
    imgs = np.stack([permute_words(words, p) for p in T])
    s = imgs.min(axis=0)
    stab = (imgs == s[None, :]).sum(axis=0)
    return s, len(T) // stab


def _host_rows(rows, basis, orbit, terms, T, x_at):
    """y[r] = sum_t [ (b_r & m_t) == R_t ] a_t sqrt(N_r / N_s) x[s] with the
    image (b_r & ~m_t) | C_t and s its representative; also sum |.| per row."""
    mask, row, col, amp = terms
    r_idx, img, a_of = [], [], []
    for k, r in enumerate(rows):
        b = int(basis[r])
        for t in range(mask.size):
            m = int(mask[t])
            if (b & m) == int(row[t]):
                r_idx.append(k)
                img.append((b & ~m) | int(col[t]))
                a_of.append(amp[t])
    r_idx = np.array(r_idx)
    img = np.array(img, np.uint64)
    a_of = np.array(a_of)
    s, n_s = _canonical(img, T)
    idx = np.searchsorted(basis, s)
    assert np.all(idx < len(basis)) and np.array_equal(basis[idx], s), "an image's representative is missing"
    n_r = orbit[np.asarray(rows)][r_idx].astype(np.float64)
    w = a_of * np.sqrt(n_r / n_s)
    xv = x_at(idx)
    y = np.zeros(len(rows))
    mag = np.zeros(len(rows))
    np.add.at(y, r_idx, w * xv)
    np.add.at(mag, r_idx, np.abs(w * xv))
    return y, mag


def test_sampled_rows_vs_host_sums(ed, sector):
    import torch

    hs, T, rh, opr, terms = sector
    assert rh.dimension == D6
    assert rh.canonicaliser() == "torus"  # 36 translations of the 6x6 torus: the compile-time 6x6 kernels
    basis = rh.basis_list
    orbit = rh.orbit_sizes
    rng = np.random.default_rng(2024)
    rows = np.sort(np.concatenate([rng.choice(D6, 960, replace=False), [0, 1, D6 - 2, D6 - 1]]))
    # every sampled word is its own representative with the stored orbit size
    s, n = _canonical(basis[rows], T)
    assert np.array_equal(s, basis[rows]) and np.array_equal(n, orbit[rows].astype(np.int64))

    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(D6, dtype=torch.float64, device="cuda", generator=g)
    x32 = x.float()

    def at(vec):
        return lambda idx: vec[torch.as_tensor(idx, device="cuda")].double().cpu().numpy()

    ref64, mag64 = _host_rows(rows, basis, orbit, terms, T, at(x))
    ref32, mag32 = _host_rows(rows, basis, orbit, terms, T, at(x32))
    rows_t = torch.as_tensor(rows, device="cuda")
    ys = {}
    # auto spreads x over the word positions in both precisions (the colex
    # form is dense enough for FP64); the blocked rank index on request
    for dt, asked in ((torch.float64, "bitmap"), (torch.float64, None), (torch.float32, None)):
        xv = x if dt == torch.float64 else x32
        if asked:
            opr.set_strategy("terms", asked)
        for transpose in (False, True):
            y = torch.empty_like(xv)
            if transpose:
                ed.mul_(y, xv, opr)
            else:
                ed.mul_(y, opr, xv)
            assert opr.last_strategy() == ("terms", asked or "xspread")
            got = y[rows_t].double().cpu().numpy()
            ref, mag = (ref64, mag64) if dt == torch.float64 else (ref32, mag32)
            tol = 1e-13 if dt == torch.float64 else 1e-5
            err = np.abs(got - ref) / mag
            assert err.max() < tol, (dt, asked, transpose, err.max())
            if not asked:
                ys[(dt, transpose)] = y
        opr.set_strategy()
    # the real symmetric matrix: row form = column form; FP32 products agree
    # with FP64 ones to single precision over the whole vector
    y64, y64c = ys[(torch.float64, False)], ys[(torch.float64, True)]
    assert (torch.linalg.norm(y64c - y64) / torch.linalg.norm(y64)).item() < 1e-14
    for tr in (False, True):
        d = torch.linalg.norm(ys[(torch.float32, tr)].double() - y64) / torch.linalg.norm(y64)
        assert d.item() < 1e-6
    del ys, y64, y64c, x, x32


def test_ground_state_vs_space_group_a1(ed, sector):
    """E0 of the translation sector (FP64 and mixed precision) against the
    A1 sector of translations x C4v (group and characters from conftest's
    generator construction), which holds the ground state: 1e-10 relative."""
    hs, T, rh, opr, terms = sector
    kw = dict(tol=1e-11, criterion="eigenvalue", max_basis=90)
    e64 = ed.eigs(opr, 1, **kw)
    emx = ed.eigs(opr, 1, mode="mixed", **kw)
    ed.release_workspace()
    perms, chars = space_group_independent(L, (0, 0), "A1")
    ops = [(ed.SitePermutation([int(j) + 1 for j in p]), complex(c)) for p, c in zip(perms, chars)]
    ra = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ops)
    assert ra.dimension == 31566122
    op, _ = _terms(ed)
    ea = ed.eigs(ed.represent(ra, op), 1, **kw)
    ed.release_workspace()
    assert e64.converged and emx.converged and ea.converged
    a, m, s = e64.values[0], emx.values[0], ea.values[0]
    assert abs(a - s) < 1e-10 * abs(s), (a, s)
    assert abs(m - s) < 1e-10 * abs(s), (m, s)
    assert abs(a - E0_REF) < 1e-8
