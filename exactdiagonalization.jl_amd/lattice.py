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
