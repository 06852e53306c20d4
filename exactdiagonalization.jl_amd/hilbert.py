"""Hilbert spaces and sectors (host-side description).

Mirrors src/HilbertSpace/site.jl, hilbert_space.jl and hilbert_space_sector.jl
of ExactDiagonalization.jl v0.14.3: a site is a list of states with quantum
numbers; site i occupies bits [bitoffsets[i], bitoffsets[i+1]) of a UInt64
word; a sector keeps the allowed total quantum numbers.  Indices are 0-based.
"""
from __future__ import annotations

import math
from typing import Iterable, Sequence

from ._lib import ArgumentError


def _qn(q):
    return tuple(q) if isinstance(q, (tuple, list)) else (int(q),)


class State:
    """State(name, quantum_number): site.jl:29-41."""

    def __init__(self, name: str, quantum_number=0):
        self.name = name
        self.quantum_number = _qn(quantum_number)

    def __eq__(self, other):
        return isinstance(other, State) and (self.name, self.quantum_number) == (other.name, other.quantum_number)

    def __hash__(self):
        return hash((self.name, self.quantum_number))

    def __repr__(self):
        return f"State({self.name!r}, {self.quantum_number})"


class Site:
    """Site(states): site.jl:69-72; bitwidth = ceil(log2(#states)) (site.jl:93)."""

    def __init__(self, states: Sequence[State]):
        if len(states) == 0:
            raise ArgumentError("a site needs at least one state")
        qd = {len(s.quantum_number) for s in states}
        if len(qd) != 1:
            raise ArgumentError("quantum numbers of a site must have the same length")
        self.states = list(states)

    @property
    def bitwidth(self) -> int:
        n = len(self.states)
        return int(math.ceil(math.log2(n))) if n > 1 else 0

    @property
    def dimension(self) -> int:
        return len(self.states)

    def __eq__(self, other):
        return isinstance(other, Site) and self.states == other.states

    def __hash__(self):
        return hash(tuple(self.states))


class HilbertSpace:
    """HilbertSpace(sites): hilbert_space.jl:25-41."""

    def __init__(self, sites: Sequence[Site]):
        self.sites = list(sites)
        self.bitwidths = [s.bitwidth for s in self.sites]
        self.bitoffsets = [0]
        for w in self.bitwidths:
            self.bitoffsets.append(self.bitoffsets[-1] + w)

    @property
    def n_sites(self) -> int:
        return len(self.sites)

    def bitwidth(self) -> int:
        return self.bitoffsets[-1]

    def get_bitmask(self, isite: int) -> int:
        """get_bitmask(hs, isite): hilbert_space.jl:107-109 (0-based site)."""
        lo, hi = self.bitoffsets[isite], self.bitoffsets[isite + 1]
        return ((1 << hi) - 1) ^ ((1 << lo) - 1)

    def qntype_len(self) -> int:
        return len(self.sites[0].states[0].quantum_number) if self.sites else 1

    def quantum_number_sectors(self):
        """quantum_number_sectors(hs): hilbert_space.jl:119-129."""
        qns = {tuple([0] * self.qntype_len())}
        for site in self.sites:
            qns = {tuple(a + b for a, b in zip(q, st.quantum_number)) for st in site.states for q in qns}
        return sorted(qns)

    def __eq__(self, other):
        return isinstance(other, HilbertSpace) and self.sites == other.sites

    def __hash__(self):
        return hash(tuple(self.sites))


class HilbertSpaceSector:
    """HilbertSpaceSector(parent[, allowed]): hilbert_space_sector.jl:13-64.

    ``allowed`` may be an int, a QN tuple, or an iterable of either; the set is
    intersected with the reachable sectors, as every reference constructor does.
    """

    def __init__(self, parent: HilbertSpace, allowed=None):
        sectors = set(parent.quantum_number_sectors())
        if allowed is None:
            want = sectors
        elif isinstance(allowed, (int, tuple)):
            want = {_qn(allowed)}
        else:
            want = {_qn(a) for a in allowed}
        self.parent = parent
        self.allowed_quantum_numbers = sectors & want

    @property
    def sites(self):
        return self.parent.sites

    def bitwidth(self):
        return self.parent.bitwidth()

    def __eq__(self, other):
        return (
            isinstance(other, HilbertSpaceSector)
            and self.parent == other.parent
            and self.allowed_quantum_numbers == other.allowed_quantum_numbers
        )


def basespace(hs):
    return hs.parent if isinstance(hs, HilbertSpaceSector) else hs


def popcount_sets(hs) -> Iterable[int] | None:
    """For sites of two states on one bit each whose states carry the same pair
    of quantum numbers, the total quantum number fixes the number of set bits:
    Q = n*q0 + p*(q1 - q0).  Returns the allowed popcounts p of a sector (or of
    the full space), or None when the space is not of that form."""
    base = basespace(hs)
    n = base.n_sites
    if n == 0:
        return None
    pairs = {tuple(st.quantum_number for st in s.states) for s in base.sites}
    if len(pairs) != 1 or any(len(s.states) != 2 for s in base.sites):
        return None
    (q0, q1), = pairs
    if isinstance(hs, HilbertSpaceSector):
        allowed = hs.allowed_quantum_numbers
    else:
        allowed = set(base.quantum_number_sectors())
    out = set()
    for Q in allowed:
        for p in range(n + 1):
            if all(n * a + p * (b - a) == c for a, b, c in zip(q0, q1, Q)):
                out.add(p)
    return sorted(out)
