"""One share of a row-sharded product (edhip_apply_rows): the rows a GPU of a
P-GPU job would compute, with the whole input vector resident, come out
bitwise as in the whole product (GPU).  Ragged and empty ranges, both forms,
every precision, device and host vectors, and the argument checks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _heis(ed, hs, bonds, w=None, d=None):
    return ed.simplify(sum((ed.heisenberg_bond(hs, i, j, 1.0 if w is None else w[k], 1.0 if d is None else d[k])
                            for k, (i, j) in enumerate(bonds)), ed.NullOperator()))


@pytest.fixture(scope="module")
def oprs(ed):
    out = {}
    hs, _ = ed.spin_half_system(24)
    j1, j2 = ed.square_bonds(4, 6)
    H = _heis(ed, hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2))
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    out["torus4x6_pi0"] = ed.represent(ed.symmetry_reduce(hsr, ed.TranslationIrrep((4, 6), (2, 0))), H)
    rng = np.random.default_rng(3)
    out["xxz4x6"] = ed.represent(hsr, _heis(ed, hs, j1, rng.uniform(0.5, 1.5, len(j1)), rng.uniform(0.8, 1.2, len(j1))))
    hs16, _ = ed.spin_half_system(16)
    H16 = _heis(ed, hs16, [(i, (i + 1) % 16) for i in range(16)] + [(i, (i + 2) % 16) for i in range(16)],
                [1.0] * 16 + [0.4] * 16)
    out["chain16_k3"] = ed.represent(ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs16, 0)),
                                                        ed.TranslationIrrep(16, 3)), H16)
    return out


def _ranges(D, P):
    cuts = [D * k // P for k in range(P + 1)]
    return list(zip(cuts[:-1], cuts[1:]))


@pytest.mark.parametrize("name", ["torus4x6_pi0", "xxz4x6", "chain16_k3"])
@pytest.mark.parametrize("transpose", [False, True])
def test_shares_reassemble_the_product_bitwise(ed, oprs, name, transpose):
    import torch

    opr = oprs[name]
    D = opr.dimension
    dtypes = [torch.complex128, torch.complex64] if opr.is_complex else [torch.float64, torch.float32]
    g = torch.Generator(device="cuda").manual_seed(7)
    for dt in dtypes:
        x = torch.randn(D, dtype=torch.float64, device="cuda", generator=g)
        if dt.is_complex:
            x = x + 1j * torch.randn(D, dtype=torch.float64, device="cuda", generator=g)
        x = x.to(dt)
        whole = torch.zeros_like(x)
        if transpose:
            ed.mul_(whole, x, opr)
        else:
            ed.mul_(whole, opr, x)
        for P in (1, 3, 8):
            parts = []
            for a, b in _ranges(D, P):
                y = torch.full((b - a,), float("nan"), dtype=dt, device="cuda")
                opr.mul_rows(y, x, a, b, transpose=transpose)
                parts.append(y)
            assert torch.equal(torch.cat(parts), whole), (dt, P)


def test_host_vectors_and_empty_range(ed, oprs):
    opr = oprs["xxz4x6"]
    D = opr.dimension
    x = np.random.default_rng(5).standard_normal(D)
    whole = opr @ x
    a, b = D // 3, D // 3 + 1000
    out = np.zeros(b - a)
    opr.mul_rows(out, x, a, b)
    assert np.array_equal(out, whole[a:b])
    empty = np.zeros(0)
    opr.mul_rows(empty, x, 5, 5)
    out32 = np.zeros(b - a, np.float32)
    opr.mul_rows(out32, x, a, b)
    full32 = np.zeros(D, np.float32)
    ed.mul_(full32, opr, x.astype(np.float32))
    assert np.array_equal(out32, full32[a:b])


def test_argument_checks(ed, oprs):
    opr = oprs["torus4x6_pi0"]
    D = opr.dimension
    x = np.zeros(D)
    with pytest.raises(ed.ArgumentError):
        opr.mul_rows(np.zeros(10), x, D - 5, D + 5)
    with pytest.raises(ed.ArgumentError):
        opr.mul_rows(np.zeros(1), x, 6, 5)
    with pytest.raises(ed.DimensionMismatch):
        opr.mul_rows(np.zeros(9), x, 0, 10)
    with pytest.raises(ed.DimensionMismatch):
        opr.mul_rows(np.zeros(10), np.zeros(D - 1), 0, 10)
    with pytest.raises(ed.ArgumentError):
        opr.mul_rows(np.zeros(10), x.astype(complex) + 1j, 0, 10)
