"""Assembled rows (kernel "assembled"): the row form's images found once and
streamed by every later product give the oracle's products (GPU).

Covered: unreduced sectors (rank arithmetic), translation sectors with unit,
real and complex characters (row-pair tori, 1D chains), space groups (cosets
of the translations, unit and +-1 characters), a dihedral group on
the generic canonicaliser, images that leave the sector (weight class 0),
add-into (apply!) products, every lane count of the streaming kernel, the
column form (which keeps the term list), an injected allocation failure of
the entries, and Lanczos ground states against the matrix-free products."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a, complex) - np.asarray(b, complex)) / np.linalg.norm(b)


def _heis(ed, hs, bonds, w=None, d=None):
    return ed.simplify(sum((ed.heisenberg_bond(hs, i, j, 1.0 if w is None else w[k], 1.0 if d is None else d[k])
                            for k, (i, j) in enumerate(bonds)), ed.NullOperator()))


def _reduced(ed, oracle, hs, n, H, ops):
    rh = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ops)
    pairs = ops.symops_and_amplitudes() if hasattr(ops, "symops_and_amplitudes") else ops
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in pairs])
    red = oracle.symmetry_reduce(oracle.sector_basis(n, 0), n, perms, [a for _, a in pairs])
    return ed.represent(rh, H), oracle.Opr(red, ed.terms_arrays(H))


def _system(ed, oracle, name):
    if name == "xxz14":  # unreduced sector, random XXZ couplings
        L = 14
        hs, _ = ed.spin_half_system(L)
        rng = np.random.default_rng(2)
        bonds = [(i, (i + 1) % L) for i in range(L)] + [(i, (i + 3) % L) for i in range(0, L, 2)]
        H = _heis(ed, hs, bonds, rng.uniform(0.5, 1.5, len(bonds)), rng.uniform(0.8, 1.2, len(bonds)))
        return ed.represent(ed.represent(ed.HilbertSpaceSector(hs, 0)), H), \
            oracle.Opr(oracle.sector_basis(L, 0), ed.terms_arrays(H))
    if name == "chain16_k3":  # complex characters
        L = 16
        hs, _ = ed.spin_half_system(L)
        H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)] + [(i, (i + 2) % L) for i in range(L)],
                  [1.0] * L + [0.4] * L)
        return _reduced(ed, oracle, hs, L, H, ed.TranslationIrrep(L, 3))
    if name.startswith("torus"):  # torus J1-J2 with a field that leaves the sector
        Lx, Ly, k = {"torus4x6_00": (4, 6, (0, 0)), "torus4x6_pi0": (4, 6, (2, 0)),
                     "torus4x4_k12": (4, 4, (1, 2))}[name]
        n = Lx * Ly
        hs, s = ed.spin_half_system(n)
        j1, j2 = ed.square_bonds(Lx, Ly)
        H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
        H = ed.simplify(H + sum((0.3 * s(i, "z") + 0.15 * s(i, "x") for i in range(n)), ed.NullOperator()))
        return _reduced(ed, oracle, hs, n, H, ed.TranslationIrrep((Lx, Ly), k))
    if name.startswith("sg4x4"):  # space group: cosets of the translations (unit and +-1 characters)
        hs, _ = ed.spin_half_system(16)
        j1, j2 = ed.square_bonds(4, 4)
        H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
        irrep = ed.SquareSpaceIrrep(4, (0, 0), "A1" if name == "sg4x4_a1" else "B1")
        return _reduced(ed, oracle, hs, 16, H, irrep)
    if name == "dihedral14":  # generic canonicaliser
        L = 14
        hs, _ = ed.spin_half_system(L)
        H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)])
        T = ed.translation_group(L)
        inv = ed.SitePermutation([((-i) % L) + 1 for i in range(L)])
        return _reduced(ed, oracle, hs, L, H, [(t, 1.0) for t in T] + [(t * inv, 1.0) for t in T])
    raise KeyError(name)


NAMES = ["xxz14", "chain16_k3", "torus4x6_00", "torus4x6_pi0", "torus4x4_k12", "sg4x4_a1", "sg4x4_b1", "dihedral14"]


@pytest.fixture(scope="module")
def systems(ed, oracle):
    return {n: _system(ed, oracle, n) for n in NAMES}


def _vec(rng, n, dtype):
    x = rng.standard_normal(n)
    if np.iscomplexobj(np.zeros(0, dtype)):
        x = x + 1j * rng.standard_normal(n)
    return x.astype(dtype)


def _tol(dtype):
    return 1e-12 if dtype in (np.float64, np.complex128) else 1e-5


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("dtype", [np.complex128, np.float64, np.complex64, np.float32])
def test_assembled_vs_oracle(ed, systems, name, dtype):
    opr, o = systems[name]
    if opr.is_complex and not np.iscomplexobj(np.zeros(0, dtype)):
        pytest.skip("complex characters need complex vectors")
    rng = np.random.default_rng(11)
    x = _vec(rng, opr.dimension, dtype)
    ref = o.apply(np.zeros(opr.dimension, complex), x.astype(complex))
    opr.set_strategy("assembled")
    try:
        n, b, _ = opr.assembled_info()
        assert n > 0 and b >= 4 * n + 8 * (opr.dimension + 1)
        out = np.zeros(opr.dimension, dtype)
        ed.mul_(out, opr, x)
        assert opr.last_strategy() == ("assembled", None)
        assert rel(out, ref) < _tol(dtype)
        # apply! adds into the output
        y0 = _vec(rng, opr.dimension, dtype)
        acc = y0.copy()
        ed.apply_(acc, opr, x)
        assert rel(acc, ref + y0.astype(complex)) < _tol(dtype)
        # the column form keeps the term-list kernel
        col = np.zeros(opr.dimension, dtype)
        ed.mul_(col, x, opr)
        assert opr.last_strategy()[0] == "terms"
        assert rel(col, o.apply(np.zeros(opr.dimension, complex), x.astype(complex), transpose=True)) < _tol(dtype)
    finally:
        opr.set_strategy()
    assert opr.assembled_info() == (0, 0, 0.0)


def test_entry_count_is_the_image_count(ed, systems):
    """One entry per matched term: the same count as the row iterators' images."""
    import ctypes

    from edhip import _lib

    for name in ("xxz14", "torus4x6_pi0"):
        opr, _ = systems[name]
        n_img = ctypes.c_int64()
        opr.ctx.check(_lib.lib().edhip_op_count_images(opr._handle, ctypes.byref(n_img)))
        opr.set_strategy("assembled")
        try:
            assert opr.assembled_info()[0] == n_img.value > 0
        finally:
            opr.set_strategy()


@pytest.mark.parametrize("lanes", [4, 8, 16, 32])
@pytest.mark.parametrize("batch", [4, 8, 16])
def test_every_lane_count(ed, systems, lanes, batch):
    from edhip import _lib

    L = _lib.lib()
    opr, o = systems["torus4x6_pi0"]
    rng = np.random.default_rng(12)
    x = rng.standard_normal(opr.dimension)
    ref = o.matvec(x).real
    opr.set_strategy("assembled")
    assert L.edhip_debug_asm_shape(lanes, batch) == 0
    try:
        out64 = opr @ x
        out32 = np.zeros(opr.dimension, np.float32)
        ed.mul_(out32, opr, x.astype(np.float32))
        assert rel(out64, ref) < 1e-12
        assert rel(out32, ref) < 1e-5
    finally:
        assert L.edhip_debug_asm_shape(0, 0) == 0
        opr.set_strategy()
    assert L.edhip_debug_asm_shape(3, 0) != 0
    assert L.edhip_debug_asm_shape(0, 5) != 0


def test_entries_allocation_failure(ed, systems):
    from edhip import _lib

    L = _lib.lib()
    opr, o = systems["xxz14"]
    x = np.random.default_rng(13).standard_normal(opr.dimension)
    assert L.edhip_debug_fail_allocs(4, 1) == 0
    try:
        with pytest.raises(MemoryError):
            opr.set_strategy("assembled")
    finally:
        L.edhip_debug_fail_allocs(0, 0)
    assert opr.assembled_info() == (0, 0, 0.0)
    y = opr @ x
    assert opr.last_strategy()[0] == "terms"
    assert rel(y, o.matvec(x).real) < 1e-12
    opr.set_strategy("assembled")
    try:
        assert rel(opr @ x, o.matvec(x).real) < 1e-12
    finally:
        opr.set_strategy()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_lanczos_on_assembled_rows(ed, systems, mode):
    opr, _ = systems["torus4x6_00"]
    e_ref = ed.eigs(opr, 2, tol=1e-12, max_basis=60).values
    opr.set_strategy("assembled")
    try:
        r = ed.eigs(opr, 2, mode=mode, tol=1e-12, max_basis=60)
        assert opr.last_strategy() == ("assembled", None)
    finally:
        opr.set_strategy()
    assert np.allclose(r.values, e_ref, rtol=0, atol=1e-10 * abs(e_ref[0]))


@pytest.mark.parametrize("strategy", ["auto", "assembled"])
def test_block_order_gives_bitwise_equal_products(ed, systems, strategy):
    """The XCD-chunked block order (edhip_debug_block_order) only reorders the
    blocks: every chunk gives bitwise the same products, on a sector large
    enough for many windows (4x6 random-bond XXZ, 2,704,156 states) and on a
    reduced one."""
    from edhip import _lib

    L = _lib.lib()
    hs, _ = ed.spin_half_system(24)
    j1, _ = ed.square_bonds(4, 6)
    rng = np.random.default_rng(3)
    H = _heis(ed, hs, j1, rng.uniform(0.5, 1.5, len(j1)), rng.uniform(0.8, 1.2, len(j1)))
    big = ed.represent(ed.represent(ed.HilbertSpaceSector(hs, 0)), H)
    try:
        for opr in (big, systems["torus4x6_pi0"][0]):
            opr.set_strategy(strategy)
            x = rng.standard_normal(opr.dimension)
            outs = []
            for chunk in (1, 2, 5, 32, 64):
                assert L.edhip_debug_block_order(chunk) == 0
                y64 = opr @ x
                y32 = np.zeros(opr.dimension, np.float32)
                ed.mul_(y32, opr, x.astype(np.float32))
                outs.append((y64, y32))
            opr.set_strategy()
            for y64, y32 in outs[1:]:
                assert np.array_equal(y64, outs[0][0]) and np.array_equal(y32, outs[0][1])
    finally:
        assert L.edhip_debug_block_order(32) == 0
    assert L.edhip_debug_block_order(65) != 0


def test_eigs_assembles_large_reduced_spaces(ed):
    """eigs assembles the rows of a reduced space over 2^22 states for the time
    of the solve (6x6 J1-J2 space-group A1 sector, 31,566,122 states, about
    9.6 GB of entries) and releases them; the eigenvalue equals the
    matrix-free solve's; an allocation failure of the entries leaves the solve
    matrix-free; an explicit assembled strategy is kept."""
    from edhip import _lib

    L = _lib.lib()
    hs, _ = ed.spin_half_system(36)
    j1, j2 = ed.square_bonds(6, 6)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    r = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ed.SquareSpaceIrrep(6, (0, 0), "A1"))
    opr = ed.represent(r, H)
    assert opr.dimension == 31566122
    kw = dict(tol=1e-10, max_basis=100, criterion="eigenvalue")
    try:
        a = ed.eigs(opr, 1, mode="mixed", **kw)
        assert a.assembled and a.converged
        assert opr.assembled_info() == (0, 0, 0.0)
        x = np.random.default_rng(1).standard_normal(opr.dimension).astype(np.float32)
        y = np.zeros(opr.dimension, np.float32)
        ed.mul_(y, opr, x)
        assert opr.last_strategy()[0] == "terms"
        m = ed.eigs(opr, 1, mode="mixed", assemble="off", **kw)
        assert not m.assembled
        assert abs(a.values[0] - m.values[0]) < 1e-10 * abs(m.values[0])
        assert abs(m.values[0] - (-18.13714754)) < 1e-7
        assert L.edhip_debug_fail_allocs(4, 1) == 0
        f = ed.eigs(opr, 1, mode="fp64", **kw)
        assert not f.assembled and abs(f.values[0] - m.values[0]) < 1e-10 * abs(m.values[0])
        assert L.edhip_debug_fail_allocs(0, 0) == 0
        opr.set_strategy("assembled")
        e = ed.eigs(opr, 1, mode="fp64", **kw)
        assert not e.assembled and opr.last_strategy() == ("assembled", None)
        assert opr.assembled_info()[0] > 0
        assert abs(e.values[0] - m.values[0]) < 1e-10 * abs(m.values[0])
    finally:
        L.edhip_debug_fail_allocs(0, 0)
        opr.set_strategy()
        ed.release_workspace()


def test_eigs_stays_matrix_free_when_entries_cannot_name_the_weights(ed):
    """A reduced space over 2^22 states whose weights outnumber what an entry
    can name (5x6 torus at k = (1, 1): 5,170,604 states, 23 column bits and
    256 classes; 12 couplings of distinct strength, one per displacement,
    times 30 character phases make 361 weights): eigs solves it matrix-free
    instead of failing, with the same eigenvalue as assemble="off"."""
    Lx, Ly = 5, 6
    n = Lx * Ly
    hs, _ = ed.spin_half_system(n)
    disp = [(1, 0), (0, 1), (1, 1), (1, -1), (2, 0), (0, 2), (2, 1), (1, 2), (2, 2), (2, -1), (-1, 2), (2, -2)]
    bonds, w = [], []
    for k, (dx, dy) in enumerate(disp):
        for y in range(Ly):
            for x in range(Lx):
                bonds.append((x + Lx * y, (x + dx) % Lx + Lx * ((y + dy) % Ly)))
                w.append(1.0 / (1 + 0.37 * k))
    H = _heis(ed, hs, bonds, w)
    r = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ed.TranslationIrrep((Lx, Ly), (1, 1)))
    opr = ed.represent(r, H)
    assert opr.dimension > 1 << 22 and opr.is_complex
    kw = dict(tol=1e-9, max_basis=80, criterion="eigenvalue")
    a = ed.eigs(opr, 1, **kw)
    b = ed.eigs(opr, 1, assemble="off", **kw)
    assert a.converged and not a.assembled and not b.assembled
    assert a.values[0] == b.values[0]
    with pytest.raises(ed.ArgumentError):
        opr.set_strategy("assembled")
    ed.release_workspace()


def test_minres_assembles_large_reduced_spaces(ed):
    """MINRES takes assembled rows for the solve as eigs does (6x6 space-group
    A1 sector): (H + 20) x = b solved with them and with the term list (an
    explicit strategy turns the assembly off) agree to the solver tolerance,
    and the residual holds for the matrix-free product."""
    import torch

    hs, _ = ed.spin_half_system(36)
    j1, j2 = ed.square_bonds(6, 6)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    r = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ed.SquareSpaceIrrep(6, (0, 0), "A1"))
    opr = ed.represent(r, H)
    g = torch.Generator(device="cuda").manual_seed(3)
    b = torch.randn(opr.dimension, dtype=torch.float64, device="cuda", generator=g)
    try:
        xa, ita, resa, asm = ed.minres(opr, b, shift=-20.0, tol=1e-11, stats=True)
        assert asm and opr.assembled_info() == (0, 0, 0.0)
        opr.set_strategy("terms")
        xt, itt, rest, asm_t = ed.minres(opr, b, shift=-20.0, tol=1e-11, stats=True)
        assert not asm_t
    finally:
        opr.set_strategy()
        ed.release_workspace()
    assert (torch.linalg.norm(xa - xt) / torch.linalg.norm(xt)).item() < 1e-9
    Ax = torch.empty_like(xa)
    ed.mul_(Ax, opr, xa)
    Ax += 20.0 * xa
    assert (torch.linalg.norm(Ax - b) / torch.linalg.norm(b)).item() < 1e-9
