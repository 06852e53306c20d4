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


class SquareSpaceIrrep:
    """One-dimensional irreps of the space group T x| C4v of an L x L torus at
    a momentum fixed by every point operation, k = (0, 0) or (L/2, L/2):
    chi(T(a, b) p) = exp(-2 pi i (kx a + ky b) / L) * chi_point(p) for the
    point irreps A1, A2, B1, B2 (the role of LatticeTools' space-symmetry
    irrep components in examples/spinhalf_triangular.jl).  8 L^2 elements,
    identity first."""

    def __init__(self, L: int, k=(0, 0), point: str = "A1"):
        if point not in _D4_IRREPS:
            raise ValueError("point irrep must be one of A1, A2, B1, B2")
        k = tuple(k)
        if any((2 * kk) % L for kk in k) or k[0] != k[1]:
            raise ValueError("the little group of k is the whole point group only for k = (0,0) or (L/2, L/2)")
        self.L, self.k, self.point = L, k, point

    def symops_and_amplitudes(self):
        L = self.L
        tr = TranslationIrrep((L, L), self.k)
        out = []
        for m in _D4:
            p = point_perm(L, m)
            cp = _D4_IRREPS[self.point][_d4_class(m)]
            for b in range(L):
                for a in range(L):
                    out.append((translation_perm((L, L), a, b) * p, tr.character(a, b) * cp))
        return out

    @property
    def group_order(self):
        return 8 * self.L * self.L


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
