"""apply!/mul! on the device vs the oracle's apply_serial! (GPU).

Tolerances: FP64 / ComplexF64 vectors 1e-12 relative (2-norm); FP32 /
ComplexF32 vectors 1e-5 relative (single-precision products and sums)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.complex128: 1e-12, np.float32: 1e-5, np.complex64: 1e-5}


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.complex128) - np.asarray(b, np.complex128)) / max(np.linalg.norm(b), 1e-300)


def _heis(ed, hs, bonds, J=None, D=None):
    ops = []
    for k, (i, j) in enumerate(bonds):
        ops.append(ed.heisenberg_bond(hs, i, j, 1.0 if J is None else J[k], 1.0 if D is None else D[k]))
    return ed.simplify(sum(ops, ed.NullOperator()))


@pytest.fixture(scope="module")
def systems(ed, oracle):
    out = {}
    # unreduced XXZ chain with random bonds (rank-indexed sector)
    L = 14
    hs, _ = ed.spin_half_system(L)
    rng = np.random.default_rng(2)
    bonds = [(i, (i + 1) % L) for i in range(L)] + [(i, (i + 3) % L) for i in range(0, L, 2)]
    H = _heis(ed, hs, bonds, rng.uniform(0.5, 1.5, len(bonds)), rng.uniform(0.8, 1.2, len(bonds)))
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    parent = oracle.sector_basis(L, 0)
    out["xxz14"] = (ed.represent(hsr, H), oracle.Opr(parent, ed.terms_arrays(H)))
    # reduced chain, complex characters (torus canonicaliser, 1D)
    L = 16
    hs, _ = ed.spin_half_system(L)
    H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)] + [(i, (i + 2) % L) for i in range(L)], [1.0] * L + [0.4] * L)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    irrep = ed.TranslationIrrep(L, 3)
    rh = ed.symmetry_reduce(hsr, irrep)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in irrep.symops_and_amplitudes()])
    red = oracle.symmetry_reduce(oracle.sector_basis(L, 0), L, perms, [a for _, a in irrep.symops_and_amplitudes()])
    out["chain16_k3"] = (ed.represent(rh, H), oracle.Opr(red, ed.terms_arrays(H)))
    # reduced 4x4 torus J1-J2 at (pi, 0): real characters
    hs, _ = ed.spin_half_system(16)
    j1, j2 = ed.square_bonds(4, 4)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    irrep = ed.TranslationIrrep((4, 4), (2, 0))
    rh = ed.symmetry_reduce(hsr, irrep)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in irrep.symops_and_amplitudes()])
    red = oracle.symmetry_reduce(oracle.sector_basis(16, 0), 16, perms, [a for _, a in irrep.symops_and_amplitudes()])
    out["sq4x4_pi0"] = (ed.represent(rh, H), oracle.Opr(red, ed.terms_arrays(H)))
    # dihedral group (generic canonicaliser), k = 0 even parity
    L = 14
    hs, _ = ed.spin_half_system(L)
    H = _heis(ed, hs, [(i, (i + 1) % L) for i in range(L)])
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    T = ed.translation_group(L)
    inv = ed.SitePermutation([((-i) % L) + 1 for i in range(L)])
    ops = [(t, 1.0) for t in T] + [(t * inv, 1.0) for t in T]
    rh = ed.symmetry_reduce(hsr, ops)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in ops])
    red = oracle.symmetry_reduce(oracle.sector_basis(L, 0), L, perms, [a for _, a in ops])
    out["dihedral14"] = (ed.represent(rh, H), oracle.Opr(red, ed.terms_arrays(H)))
    return out


@pytest.mark.parametrize("name", ["xxz14", "chain16_k3", "sq4x4_pi0", "dihedral14"])
@pytest.mark.parametrize("dtype", [np.complex128, np.float64, np.complex64, np.float32])
@pytest.mark.parametrize("transpose", [False, True])
def test_apply_vs_oracle(ed, systems, name, dtype, transpose):
    opr, o = systems[name]
    if opr.is_complex and not np.iscomplexobj(np.zeros(0, dtype)):
        with pytest.raises(ed.ArgumentError):
            ed.mul_(np.zeros(opr.dimension, dtype), opr, np.zeros(opr.dimension, dtype))
        return
    rng = np.random.default_rng(4)
    x = rng.standard_normal(opr.dimension)
    if np.iscomplexobj(np.zeros(0, dtype)):
        x = x + 1j * rng.standard_normal(opr.dimension)
    x = x.astype(dtype)
    y0 = rng.standard_normal(opr.dimension).astype(dtype)
    ref = o.apply(y0.astype(np.complex128), x.astype(np.complex128), transpose=transpose)
    out = y0.copy()
    if transpose:
        ed.apply_(out, x, opr)  # apply!(out, state, opr): out += state * opr
    else:
        ed.apply_(out, opr, x)  # apply!(out, opr, state): out += opr * state
    assert rel(out, ref) < TOL[dtype]
    # mul! overwrites
    out2 = y0.copy()
    ed.mul_(out2, *((x, opr) if transpose else (opr, x)))
    assert rel(out2, ref - y0) < TOL[dtype] * 10


def test_reference_apply_semantics(ed, golden):
    """test_operator_representation.jl:181-272: apply! adds, mul! overwrites,
    row-vector form, promotion, DimensionMismatch."""
    hs, s = ed.spin_half_system(4)
    hsr = ed.represent(hs)
    dim = hsr.dimension
    opr = ed.represent(hsr, s(1, "+"))
    dense = np.array(golden["known_answers"]["sigma_plus_site2_n4"]["dense"], float)
    rng = np.random.default_rng(0)
    state = rng.random(dim) + 1j * rng.random(dim)
    out1 = np.zeros(dim, complex)
    ed.apply_(out1, opr, state)
    assert np.allclose(out1, dense @ state, atol=1e-6)
    assert np.allclose(opr @ state, dense @ state, atol=1e-6)
    ed.apply_(out1, opr, state)
    assert not np.allclose(out1, dense @ state, atol=1e-6)
    out1[:] = 0
    ed.apply_(out1, state, opr)
    assert np.allclose(out1, state @ dense, atol=1e-6)
    assert np.allclose(state @ opr, state @ dense, atol=1e-6)
    with pytest.raises(ed.DimensionMismatch):
        ed.apply_(out1, opr, np.zeros(dim + 1, complex))
    with pytest.raises(ed.DimensionMismatch):
        ed.apply_(np.zeros(dim + 1, complex), opr, state)
    with pytest.raises(ed.DimensionMismatch):
        ed.apply_(out1, np.zeros(dim + 1, complex), opr)
    # mul! overwrites, apply! does not
    ox = ed.represent(hsr, s(1, "x"))
    vin = np.arange(1, dim + 1) * 0.1
    a0, a1 = np.zeros(dim), np.arange(1, dim + 1) * 10.0
    ed.apply_(a0, ox, vin)
    ed.apply_(a1, ox, vin)
    assert not np.allclose(a0, a1, atol=1e-6)
    m0, m1 = np.zeros(dim), np.arange(1, dim + 1) * 10.0
    ed.mul_(m0, ox, vin)
    ed.mul_(m1, ox, vin)
    assert np.allclose(m0, m1, atol=1e-6)
    # promotion: complex operator on a real vector gives a complex result
    oy = ed.represent(hsr, s(1, "y"))
    assert (oy @ np.ones(dim)).dtype == np.complex128
    assert (ox @ np.ones(dim)).dtype == np.float64


def test_images_outside_sector_are_dropped(ed):
    """test_operator_representation.jl:275-296: sigma^x changes Sz, so every
    image leaves the Sz = 0 sector and apply! adds nothing."""
    hs, s = ed.spin_half_system(4)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    opr = ed.represent(hsr, s(1, "x"))
    state = 2j * np.ones(hsr.dimension)
    out = np.zeros(hsr.dimension, complex)
    ed.apply_(out, opr, state)
    assert np.all(out == 0)
    ed.apply_(out, state, opr)
    assert np.all(out == 0)


def test_device_tensors_match_host(ed, systems):
    import torch

    opr, o = systems["chain16_k3"]
    rng = np.random.default_rng(9)
    x = rng.standard_normal(opr.dimension) + 1j * rng.standard_normal(opr.dimension)
    host = opr @ x
    xt = torch.from_numpy(x).cuda()
    yt = torch.zeros_like(xt)
    ed.apply_(yt, opr, xt)
    assert rel(yt.cpu().numpy(), host) < 1e-14


def test_device_tensors_stream_handoff(ed):
    """Torch work queued right before and read right after a device apply!
    (no explicit synchronisation by the caller) sees the ordered result: a
    2,704,156-state matvec whose input is produced by a torch kernel and whose
    output is consumed by one."""
    import torch

    n = 24
    hs, _ = ed.spin_half_system(n)
    H = ed.simplify(sum((ed.heisenberg_bond(hs, i, (i + 1) % n, 1.0) for i in range(n)), ed.NullOperator()))
    opr = ed.represent(ed.represent(ed.HilbertSpaceSector(hs, 0)), H)
    ref = float(np.sum(opr @ np.ones(opr.dimension)))
    g = torch.Generator(device="cuda").manual_seed(5)
    for rep in range(3):
        x = torch.randn(opr.dimension, dtype=torch.float64, device="cuda", generator=g) * 0 + 1.0
        y = torch.full_like(x, float(rep + 7))
        ed.mul_(y, opr, x)
        s = float(y.sum().item())
        assert abs(s - ref) < 1e-9 * max(1.0, abs(ref))


def test_device_tensors_side_stream(ed, systems):
    """The same hand-off on a caller's non-default torch stream (the library
    adopts it as its own queue), then back on the default stream: products
    chained through torch kernels without host synchronisation agree with
    the host path."""
    import torch

    opr, _ = systems["chain16_k3"]
    rng = np.random.default_rng(4)
    x = rng.standard_normal(opr.dimension) + 1j * rng.standard_normal(opr.dimension)
    ref = opr @ (opr @ (2.0 * x))
    side = torch.cuda.Stream()
    for stream in (side, torch.cuda.current_stream()):
        with torch.cuda.stream(stream):
            xt = torch.from_numpy(x).to("cuda", non_blocking=False) * 2.0
            y1 = torch.zeros_like(xt)
            ed.mul_(y1, opr, xt)
            y2 = torch.zeros_like(xt)
            ed.mul_(y2, opr, y1 * 1.0)
            out = y2.cpu().numpy()
        assert rel(out, ref) < 1e-13


def test_hermiticity_and_linearity(ed, systems):
    for name in ("chain16_k3", "sq4x4_pi0", "dihedral14", "xxz14"):
        opr, _ = systems[name]
        rng = np.random.default_rng(1)
        d = opr.dimension
        x = rng.standard_normal(d) + 1j * rng.standard_normal(d)
        y = rng.standard_normal(d) + 1j * rng.standard_normal(d)
        Hx, Hy = opr @ x, opr @ y
        assert abs(np.vdot(y, Hx) - np.vdot(Hy, x)) < 1e-11 * np.linalg.norm(x) * np.linalg.norm(y) * 10
        assert rel(opr @ (2.0 * x - 3.0j * y), 2.0 * Hx - 3.0j * Hy) < 1e-13


@pytest.mark.parametrize("n,p", [(20, 3), (30, 2), (40, 2), (48, 2), (63, 2)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_wide_words_few_particles(ed, oracle, n, p, dtype):
    """Sectors of few up spins on wide words (K = 20 ... 63 bits, the widest
    a UInt64 space allows): the spread
    of x takes its colex form where the positions fit in 32 bits, the plain
    form or none otherwise (then the rank index or the sorted search); an
    XY ring with next-nearest hopping and random fields (every image keeps
    the popcount: the products skip the popcount test) and the same plus a
    transverse field on two sites (images of another popcount, dropped), in
    the row and column forms, against the oracle."""
    hs, s = ed.spin_half_system(n)
    rng = np.random.default_rng(n + p)
    hop = [(i, (i + 1) % n) for i in range(n)] + [(i, (i + 3) % n) for i in range(n)]
    H = ed.simplify(sum([ed.heisenberg_bond(hs, i, j, rng.uniform(0.5, 1.5), 0.0) for i, j in hop]
                        + [rng.uniform(-1, 1) * s(i, "z") for i in range(n)], ed.NullOperator()))
    Hx = ed.simplify(H + 0.7 * s(1, "x") + 0.4 * s(n - 2, "x"))
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 2 * p - n))
    basis = np.array([int(b) for b in hsr.basis_list], dtype=np.uint64)
    assert basis.size == math.comb(n, p)
    x = rng.standard_normal(basis.size).astype(dtype)
    for op in (H, Hx):
        opr = ed.represent(hsr, op)
        o = oracle.Opr(basis, ed.terms_arrays(op))
        for transpose in (False, True):
            ref = o.apply(np.zeros(basis.size, complex), x.astype(complex), transpose=transpose)
            out = np.zeros(basis.size, dtype)
            if transpose:
                ed.mul_(out, x, opr)
            else:
                ed.mul_(out, opr, x)
            assert rel(out, ref) < TOL[dtype], (n, p, transpose)
