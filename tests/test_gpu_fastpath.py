"""Every kernel / lookup combination (term list or bond pairs; sorted search,
bitmap rank index or hash table) gives the same products as the oracle (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a, complex) - np.asarray(b, complex)) / np.linalg.norm(b)


def _heis(ed, hs, bonds, w=None, d=None):
    return ed.simplify(sum((ed.heisenberg_bond(hs, i, j, 1.0 if w is None else w[k], 1.0 if d is None else d[k])
                            for k, (i, j) in enumerate(bonds)), ed.NullOperator()))


def _torus_system(ed, oracle, Lx, Ly, k, j2=0.5, field=0.0):
    n = Lx * Ly
    hs, s = ed.spin_half_system(n)
    j1, jj2 = ed.square_bonds(Lx, Ly)
    H = _heis(ed, hs, j1 + jj2, [1.0] * len(j1) + [j2] * len(jj2))
    if field:
        # one-bit terms: a longitudinal field (diagonal) and a transverse one
        # (images leave the Sz sector and are dropped)
        H = ed.simplify(H + sum((field * s(i, "z") + 0.5 * field * s(i, "x") for i in range(n)), ed.NullOperator()))
    irrep = ed.TranslationIrrep((Lx, Ly), k)
    rh = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), irrep)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in irrep.symops_and_amplitudes()])
    red = oracle.symmetry_reduce(oracle.sector_basis(n, 0), n, perms, [a for _, a in irrep.symops_and_amplitudes()])
    return ed.represent(rh, H), oracle.Opr(red, ed.terms_arrays(H))


STRATS = [("terms", "search"), ("terms", "bitmap"), ("terms", "xspread"), ("pairs", "search"), ("pairs", "bitmap"),
          ("pairs", "hash")]


@pytest.mark.parametrize("Lx,Ly,k", [(4, 4, (2, 0)), (4, 4, (1, 2)), (4, 6, (2, 0)), (4, 6, (1, 3)), (6, 4, (3, 0)),
                                     (6, 4, (1, 2)), (3, 4, (0, 1)), (5, 4, (2, 1)),
                                     # k = (0, 0): the unit-phase translation-only kernels
                                     (4, 4, (0, 0)), (4, 6, (0, 0)), (6, 4, (0, 0))])
@pytest.mark.parametrize("dtype", [np.complex128, np.float64, np.float32, np.complex64])
@pytest.mark.parametrize("transpose", [False, True])
def test_strategies_vs_oracle(ed, oracle, Lx, Ly, k, dtype, transpose):
    opr, o = _torus_system(ed, oracle, Lx, Ly, k, field=0.3)
    if opr.is_complex and not np.iscomplexobj(np.zeros(0, dtype)):
        pytest.skip("complex characters need complex vectors")
    rng = np.random.default_rng(8)
    x = rng.standard_normal(opr.dimension)
    if np.iscomplexobj(np.zeros(0, dtype)):
        x = x + 1j * rng.standard_normal(opr.dimension)
    x = x.astype(dtype)
    ref = o.apply(np.zeros(opr.dimension, complex), x.astype(complex), transpose=transpose)
    tol = 1e-12 if np.dtype(dtype).itemsize >= 8 and dtype != np.complex64 else 1e-5
    for kernel, lookup in STRATS:
        opr.set_strategy(kernel, lookup)
        out = np.zeros(opr.dimension, dtype)
        if transpose:
            ed.mul_(out, x, opr)
        else:
            ed.mul_(out, opr, x)
        assert opr.last_strategy() == (kernel, lookup)
        assert rel(out, ref) < tol, (kernel, lookup)
    opr.set_strategy()


def test_unreduced_pairs_rank(ed, oracle):
    """Unreduced sector (rank arithmetic) with the pair kernel: XXZ bonds."""
    L = 16
    hs, s = ed.spin_half_system(L)
    rng = np.random.default_rng(3)
    bonds = [(i, (i + 1) % L) for i in range(L)] + [(i, (i + 5) % L) for i in range(L)]
    H = _heis(ed, hs, bonds, rng.uniform(0.5, 1.5, len(bonds)), rng.uniform(0.8, 1.2, len(bonds)))
    opr = ed.represent(ed.represent(ed.HilbertSpaceSector(hs, 0)), H)
    o = oracle.Opr(oracle.sector_basis(L, 0), ed.terms_arrays(H))
    x = rng.standard_normal(opr.dimension)
    ref = o.matvec(x).real
    for kernel in ("terms", "pairs"):
        opr.set_strategy(kernel)
        assert rel(opr @ x, ref) < 1e-13
    assert opr.last_strategy() == ("pairs", "search")


def test_pairs_rejected_where_not_covered(ed):
    L = 14
    hs, s = ed.spin_half_system(L)
    H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)])
    rh = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ed.TranslationIrrep(L, 1))
    opr = ed.represent(rh, H)  # ring of 14: rows wider than the 12-bit rotation table
    with pytest.raises(ed.ArgumentError):
        opr.set_strategy("pairs")
    three = ed.represent(rh, ed.simplify(s(0, "z") * s(1, "z") * s(2, "z")))
    with pytest.raises(ed.ArgumentError):
        three.set_strategy("pairs")


def test_hash_table_widening(ed, oracle):
    """The table is widened from 16- to 32-byte slots when a ComplexF64 product
    follows a real one; real products keep working on the wide table."""
    opr, o = _torus_system(ed, oracle, 4, 4, (2, 0))
    opr.set_strategy("pairs", "hash")
    rng = np.random.default_rng(1)
    xr = rng.standard_normal(opr.dimension)
    y1 = opr @ xr
    xc = xr + 1j * rng.standard_normal(opr.dimension)
    y2 = opr @ xc
    y3 = opr @ xr
    assert rel(y1, o.matvec(xr)) < 1e-13 and rel(y2, o.matvec(xc)) < 1e-13 and rel(y3, y1) < 1e-15
    y4 = (opr @ xr.astype(np.float32)).astype(np.float64)
    assert rel(y4, y1) < 1e-5


def test_torus6x6_properties(ed):
    """The 6x6 J1-J2 k=(0,0) sector (252,091,362 representatives), where no
    oracle fits: the product is Hermitian, the row and column forms agree, the
    bond-pair kernel agrees with the term-list kernel, and the FP32 product
    agrees with the FP64 one to FP32 accuracy."""
    import torch

    Lx = Ly = 6
    hs, _ = ed.spin_half_system(36)
    j1, j2 = ed.square_bonds(Lx, Ly)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    rh = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ed.TranslationIrrep((Lx, Ly), (0, 0)))
    opr = ed.represent(rh, H)
    D = opr.dimension
    assert D == 252091362
    g = torch.Generator(device="cuda").manual_seed(11)
    u = torch.randn(D, dtype=torch.float64, device="cuda", generator=g)
    v = torch.randn(D, dtype=torch.float64, device="cuda", generator=g)
    Hv = torch.empty_like(v)
    Hu = torch.empty_like(u)
    ed.mul_(Hv, opr, v)
    ed.mul_(Hu, opr, u)
    a, b = torch.dot(u, Hv).item(), torch.dot(Hu, v).item()
    assert abs(a - b) <= 1e-11 * abs(a)
    vH = torch.empty_like(v)
    ed.mul_(vH, v, opr)
    assert (torch.linalg.norm(vH - Hv) / torch.linalg.norm(Hv)).item() < 1e-13
    assert opr.last_strategy() == ("terms", "xspread")  # FP64: the dense (colex) spread of x
    for strat in (("pairs", "bitmap"), ("terms", "bitmap")):
        opr.set_strategy(*strat)
        Hv2 = torch.empty_like(v)
        ed.mul_(Hv2, opr, v)
        assert opr.last_strategy() == strat
        assert (torch.linalg.norm(Hv2 - Hv) / torch.linalg.norm(Hv)).item() < 1e-13
    opr.set_strategy()
    del Hu, vH, Hv2, u
    v32 = v.float()
    Hv32 = torch.empty_like(v32)
    ed.mul_(Hv32, opr, v32)
    assert opr.last_strategy() == ("terms", "xspread")  # FP32: x spread over the word positions
    assert (torch.linalg.norm(Hv32.double() - Hv) / torch.linalg.norm(Hv)).item() < 1e-5
    opr.set_strategy("terms", "bitmap")
    Hv32b = torch.empty_like(v32)
    ed.mul_(Hv32b, opr, v32)
    opr.set_strategy()
    # same images in the same order; the spread holds x / sqrt(N) rounded to
    # FP32 where the rank path rounds the weight a sqrt(N_r / N_s): equal to
    # FP32 rounding
    assert (torch.linalg.norm(Hv32 - Hv32b) / torch.linalg.norm(Hv32b)).item() < 2e-7
