"""Device allocation failures, injected from the host
(edhip_debug_fail_allocs), take the documented fallbacks (GPU):

* the spread of x over the word positions: the products read x through the
  rank index instead, with the same results;
* the eigensolver's Krylov workspace: the spreads are freed and the
  allocation tried once more; a second failure raises MemoryError and
  leaves the context usable."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a, complex) - np.asarray(b, complex)) / np.linalg.norm(b)


@pytest.fixture
def inject():
    from edhip import _lib

    L = _lib.lib()

    def set_(kinds, count):
        assert L.edhip_debug_fail_allocs(kinds, count) == 0

    yield set_
    set_(0, 0)


def _mul32(ed, opr, x):
    """An FP32 product (opr @ x would promote to the operator's FP64)."""
    out = np.zeros(opr.dimension, np.float32)
    ed.mul_(out, opr, x.astype(np.float32))
    return out


def _torus(ed, oracle):
    Lx, Ly = 4, 6
    n = Lx * Ly
    hs, s = ed.spin_half_system(n)
    j1, j2 = ed.square_bonds(Lx, Ly)
    H = ed.simplify(sum([ed.heisenberg_bond(hs, i, j, 1.0) for i, j in j1]
                        + [ed.heisenberg_bond(hs, i, j, 0.5) for i, j in j2], ed.NullOperator()))
    irrep = ed.TranslationIrrep((Lx, Ly), (0, 0))
    rh = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), irrep)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in irrep.symops_and_amplitudes()])
    red = oracle.symmetry_reduce(oracle.sector_basis(n, 0), n, perms, [a for _, a in irrep.symops_and_amplitudes()])
    return ed.represent(rh, H), oracle.Opr(red, ed.terms_arrays(H))


def test_spread_allocation_failure_falls_back(ed, oracle, inject):
    opr, o = _torus(ed, oracle)
    x = np.random.default_rng(4).standard_normal(opr.dimension)
    ref = o.matvec(x).real
    inject(1, 1)  # the first spread of x fails
    y32 = _mul32(ed, opr, x)
    assert opr.last_strategy() == ("terms", "bitmap")
    assert rel(y32, ref) < 1e-5
    y64 = opr @ x  # the FP64 spread is a separate allocation: it succeeds
    assert opr.last_strategy() == ("terms", "xspread")
    assert rel(y64, ref) < 1e-13
    inject(0, 0)
    y32b = _mul32(ed, opr, x)  # the failed spread is not tried again
    assert opr.last_strategy() == ("terms", "bitmap")
    assert np.array_equal(y32, y32b)


def test_workspace_allocation_failure_retries_then_raises(ed, oracle, inject):
    opr, o = _torus(ed, oracle)
    x = np.random.default_rng(5).standard_normal(opr.dimension)
    _mul32(ed, opr, x)  # builds the FP32 spread
    assert opr.last_strategy() == ("terms", "xspread")
    e_ref = ed.eigs(opr, 1, tol=1e-12, max_basis=60).values[0]
    ed.release_workspace()
    inject(2, 1)  # the first workspace allocation fails: freed spreads, a second try succeeds
    r = ed.eigs(opr, 1, mode="mixed", tol=1e-12, max_basis=60)
    assert abs(r.values[0] - e_ref) < 1e-10 * abs(e_ref)
    ed.release_workspace()
    inject(2, 2)  # both tries fail
    with pytest.raises(MemoryError):
        ed.eigs(opr, 1, tol=1e-12, max_basis=60)
    inject(0, 0)
    r2 = ed.eigs(opr, 1, tol=1e-12, max_basis=60)
    assert abs(r2.values[0] - e_ref) < 1e-12 * abs(e_ref)
    y = _mul32(ed, opr, x)  # the spread is rebuilt after being freed
    assert opr.last_strategy() == ("terms", "xspread")
    assert rel(y, o.matvec(x).real) < 1e-5
