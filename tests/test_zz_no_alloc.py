"""edhip_term_expect_noalloc (the per-term forms with the partial sums in a
fixed per-GPU scratch array, no device allocation) against edhip_term_expect,
through the C ABI: two calls in a row, the second on a smaller operator, so
that the second reads scratch the first has filled (GPU).  Tolerance: 1e-13
of the largest form (the same kernels and fixed-order reductions; only the
way psi is fetched differs)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _xxz(ed, bonds, n, seed):
    hs, _ = ed.spin_half_system(n)
    rng = np.random.default_rng(seed)
    return hs, ed.simplify(sum((ed.heisenberg_bond(hs, i, j, rng.uniform(0.5, 1.5), rng.uniform(0.8, 1.2))
                                for i, j in bonds), ed.NullOperator()))


def _both(ed, opr, prec, phi, psi):
    from edhip import _lib

    L = _lib.lib()
    nt = opr._n_terms
    a = np.full(2 * nt, np.nan)
    b = np.full(2 * nt, np.nan)
    opr.ctx.synchronize()
    import torch

    torch.cuda.synchronize()
    for fn, out in ((L.edhip_term_expect, a), (L.edhip_term_expect_noalloc, b)):
        opr.ctx.check(fn(opr._handle, prec, ctypes.c_void_p(phi.data_ptr()), ctypes.c_void_p(psi.data_ptr()),
                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return a, b


def test_noalloc_matches_term_expect(ed):
    import torch
    from edhip import _lib

    # first call: 4x4 torus XXZ (J1 + J2 bonds: 64 bonds, 384 terms) on the
    # k = (1, 0) translation sector, complex vectors
    j1, j2 = ed.square_bonds(4, 4)
    hs, H = _xxz(ed, j1 + j2, 16, 1)
    rh = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), ed.TranslationIrrep((4, 4), (1, 0)))
    op1 = ed.represent(rh, H)
    g = torch.Generator(device="cuda").manual_seed(2)
    d1 = op1.dimension
    phi = torch.randn(d1, dtype=torch.complex128, device="cuda", generator=g)
    psi = torch.randn(d1, dtype=torch.complex128, device="cuda", generator=g)
    a, b = _both(ed, op1, _lib.C128, phi, psi)
    assert op1._n_terms <= 512 and not np.isnan(b).any()
    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
    # second call: a smaller operator (10-site chain, unreduced sector, FP64)
    hs2, H2 = _xxz(ed, [(i, (i + 1) % 10) for i in range(10)], 10, 3)
    op2 = ed.represent(ed.represent(ed.HilbertSpaceSector(hs2, 0)), H2)
    d2 = op2.dimension
    phi2 = torch.randn(d2, dtype=torch.float64, device="cuda", generator=g)
    psi2 = torch.randn(d2, dtype=torch.float64, device="cuda", generator=g)
    a2, b2 = _both(ed, op2, _lib.F64, phi2, psi2)
    assert op2._n_terms < op1._n_terms and d2 < d1
    assert np.abs(a2 - b2).max() <= 1e-13 * np.abs(a2).max()
    # the forms are the pieces of <phi|H|psi>: their amplitude-weighted sum
    # equals the product computed by apply!
    mask, row, col, amp = ed.terms_arrays(H2)
    y = torch.empty_like(psi2)
    ed.mul_(y, op2, psi2)
    ref = torch.dot(phi2, y).item()
    assert abs(np.dot(amp.real, b2[0::2]) - ref) <= 1e-12 * max(1.0, abs(ref))


def test_noalloc_rejects_too_many_terms(ed):
    import torch
    from edhip import _lib

    L = 14
    hs, s = ed.spin_half_system(L)
    # two exchange terms per triple of sites i < j < k: 2 C(14, 3) = 728 > 512
    trip = [(i, j, k) for i in range(L) for j in range(i + 1, L) for k in range(j + 1, L)]
    op = ed.SumOperator([ed.PureOperator((1 << i) | (1 << j) | (1 << k), r, c, 0.1)
                         for i, j, k in trip for r, c in ((1 << i, 1 << k), (1 << k, 1 << i))])
    opr = ed.represent(ed.represent(ed.HilbertSpaceSector(hs, 0)), op)
    assert opr._n_terms > 512
    v = torch.zeros(opr.dimension, dtype=torch.float64, device="cuda")
    out = np.zeros(2 * opr._n_terms)
    rc = _lib.lib().edhip_term_expect_noalloc(opr._handle, _lib.F64, ctypes.c_void_p(v.data_ptr()),
                                              ctypes.c_void_p(v.data_ptr()),
                                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == _lib.EDHIP_ERR_ARGUMENT
