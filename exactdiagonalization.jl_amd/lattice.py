"""Translation groups of periodic chains and tori, and their 1D irreps.

The reference takes its symmetry groups from LatticeTools (not vendored);
this module supplies the translation part of that input in the conventions the
reference's own tests pin (test/test_symmetry_reduce.jl:164-214):
  * a site i (0-based) of an Lx x Ly torus sits at (x, y) = (i % Lx, i // Lx)
    (site = x + Lx*y, the sub2ind order of examples/spinhalfsquare_small.jl);
  * T(a, b) sends (x, y) to (x + a, y + b);
  * the character of momentum (kx, ky) is exp(-2 pi i (kx a / Lx + ky b / Ly)).
Elements are listed with the identity first, a running fastest.
"""
from __future__ import annotations

import cmath
import math

from .representation import SitePermutation


def translation_perm(shape, a: int, b: int = 0) -> SitePermutation:
    Lx, Ly = (shape, 1) if isinstance(shape, int) else shape
    m = []
    for i in range(Lx * Ly):
        x, y = i % Lx, i // Lx
        m.append(((x + a) % Lx) + Lx * ((y + b) % Ly) + 1)
    return SitePermutation(m)


def translation_group(shape):
    """All translations T(a, b), identity first."""
    Lx, Ly = (shape, 1) if isinstance(shape, int) else shape
    return [translation_perm((Lx, Ly), a, b) for b in range(Ly) for a in range(Lx)]


class TranslationIrrep:
    """The 1D irrep of momentum k of the translation group of a chain/torus
    (the role of IrrepComponent(tsymbed, k+1) in the reference tests)."""

    def __init__(self, shape, k):
        self.shape = (shape, 1) if isinstance(shape, int) else tuple(shape)
        self.k = (k, 0) if isinstance(k, int) else tuple(k)

    def character(self, a: int, b: int = 0) -> complex:
        Lx, Ly = self.shape
        kx, ky = self.k
        phase = kx * a / Lx + ky * b / Ly
        return cmath.exp(-2j * math.pi * phase)

    def symops_and_amplitudes(self):
        Lx, Ly = self.shape
        return [
            (translation_perm((Lx, Ly), a, b), self.character(a, b)) for b in range(Ly) for a in range(Lx)
        ]

    @property
    def group_order(self):
        return self.shape[0] * self.shape[1]


# The point group C4v of the square lattice about site 0: linear parts (2x2
# integer matrices acting on (x, y)) and their one-dimensional irreps.
_D4 = [((1, 0), (0, 1)), ((0, -1), (1, 0)), ((-1, 0), (0, -1)), ((0, 1), (-1, 0)),
       ((-1, 0), (0, 1)), ((1, 0), (0, -1)), ((0, 1), (1, 0)), ((0, -1), (-1, 0))]


def _d4_class(m):
    """'E', 'C2', 'C4', 'sv' (axial mirror) or 'sd' (diagonal mirror)."""
    (a, b), (c, d) = m
    if m == ((1, 0), (0, 1)):
        return "E"
    if m == ((-1, 0), (0, -1)):
        return "C2"
    if a * d - b * c == 1:
        return "C4"
    return "sv" if b == 0 else "sd"


_D4_IRREPS = {  # characters on the classes E, C2, C4, sv, sd
    "A1": {"E": 1, "C2": 1, "C4": 1, "sv": 1, "sd": 1},
    "A2": {"E": 1, "C2": 1, "C4": 1, "sv": -1, "sd": -1},
    "B1": {"E": 1, "C2": 1, "C4": -1, "sv": 1, "sd": -1},
    "B2": {"E": 1, "C2": 1, "C4": -1, "sv": -1, "sd": 1},
}


def point_perm(L: int, m) -> SitePermutation:
    """Site permutation of the linear map m about site 0 on an L x L torus."""
    (a, b), (c, d) = m
    out = []
    for i in range(L * L):
        x, y = i % L, i // L
        out.append(((a * x + b * y) % L) + L * ((c * x + d * y) % L) + 1)
    return SitePermutation(out)


# One-dimensional irreps of the little groups of C4v on a torus, by name.
# C4v: characters on its classes.  C2v (at (L/2, 0) and (0, L/2): C2 and the
# two axial mirrors; a diagonal C2v would need 2k = 0 on a diagonal, where
# the little group is all of C4v): characters on (E, C2, x -> -x, y -> -y).
# Cs (one mirror): on (E, the mirror).  C1: the identity alone.
_MX, _MY = ((-1, 0), (0, 1)), ((1, 0), (0, -1))


def _little_irreps(kind):
    if kind == "C4v":
        return dict(_D4_IRREPS)
    if kind == "C2v":
        # characters on (E, C2, first mirror, second mirror)
        return {"A1": (1, 1, 1, 1), "A2": (1, 1, -1, -1), "B1": (1, -1, 1, -1), "B2": (1, -1, -1, 1)}
    if kind == "Cs":
        return {"A'": (1, 1), "A''": (1, -1)}
    return {"A": (1,)}


class SquareSpaceIrrep:
    """One-dimensional irreps of the symmorphic space group T x| C4v of an
    L x L torus at momentum k (integers, k_i in [0, L)): the elements T(a, b) p
    with p in the little group of k (the point operations with p k = k mod L)
    and chi(T(a, b) p) = exp(-2 pi i (kx a + ky b) / L) chi_point(p), chi_point
    a one-dimensional irrep of the little group (the elements the reference's
    symmorphic reduction keeps: those with a nonzero character,
    symmetry_reduce_symmorphic.jl:27-46; LatticeTools' space-symmetry irrep
    components in examples/spinhalf_triangular.jl).  Little groups and their
    irreps: C4v at k = (0, 0) and (L/2, L/2) (A1, A2, B1, B2); C2v at (L/2, 0)
    and (0, L/2) (A1, A2, B1, B2 with the characters of _little_irreps); Cs on
    a mirror line (A', A''); C1 elsewhere (A).  |little group| L^2 elements,
    identity first."""

    def __init__(self, L: int, k=(0, 0), point: str = "A1"):
        k = (int(k[0]) % L, int(k[1]) % L)
        self.L, self.k, self.point = L, k, point
        self.little = [m for m in _D4 if self._fixes(m)]
        self.kind = {8: "C4v", 4: "C2v", 2: "Cs", 1: "C1"}[len(self.little)]
        assert self.kind != "C2v" or (_MX in self.little and _MY in self.little)
        irreps = _little_irreps(self.kind)
        if point not in irreps:
            raise ValueError(f"the little group of k = {k} is {self.kind}; its irreps are {sorted(irreps)}")

    def _fixes(self, m):
        (a, b), (c, d) = m
        kx, ky = self.k
        L = self.L
        return ((a * kx + b * ky - kx) % L, (c * kx + d * ky - ky) % L) == (0, 0)

    def point_character(self, m) -> int:
        """chi_point of a little-group element (a 2x2 integer matrix)."""
        chars = _little_irreps(self.kind)[self.point]
        if self.kind == "C4v":
            return chars[_d4_class(m)]
        if m == _D4[0]:
            return chars[0]
        if self.kind == "C2v":
            return chars[{((-1, 0), (0, -1)): 1, _MX: 2, _MY: 3}[m]]
        return chars[1]  # Cs: the mirror

    def symops_and_amplitudes(self):
        L = self.L
        tr = TranslationIrrep((L, L), self.k)
        out = []
        for m in self.little:
            p = point_perm(L, m)
            cp = self.point_character(m)
            for b in range(L):
                for a in range(L):
                    out.append((translation_perm((L, L), a, b) * p, tr.character(a, b) * cp))
        return out

    @property
    def group_order(self):
        return len(self.little) * self.L * self.L


def chain_bonds(L: int):
    """Nearest-neighbour bonds (i, i+1 mod L) of a periodic chain."""
    return [(i, (i + 1) % L) for i in range(L)]


def square_bonds(Lx: int, Ly: int):
    """(J1, J2) bonds of a periodic Lx x Ly square lattice: J1 along +x and +y,
    J2 along both diagonals (x+1, y+1) and (x+1, y-1)."""
    s = lambda x, y: (x % Lx) + Lx * (y % Ly)
    j1, j2 = [], []
    for y in range(Ly):
        for x in range(Lx):
            j1.append((s(x, y), s(x + 1, y)))
            j1.append((s(x, y), s(x, y + 1)))
            j2.append((s(x, y), s(x + 1, y + 1)))
            j2.append((s(x, y), s(x + 1, y - 1)))
    return j1, j2
