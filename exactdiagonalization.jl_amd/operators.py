"""Operators: the PureOperator / SumOperator term format and its algebra.

Host-side mirror of src/Operator/ (pure_operator.jl, sum_operator.jl,
null_operator.jl, operator_simplify.jl) and of the Pauli constructors of
src/Toolkit/spin_half.jl.  This is the pre-pass that fixes the term list the
device kernels consume: (bitmask, bitrow, bitcol, amplitude) per term.
"""
from __future__ import annotations

import math
from numbers import Number

import numpy as np

from ._lib import ArgumentError
from .hilbert import HilbertSpace, HilbertSpaceSector, basespace

RTOL_DEFAULT = math.sqrt(np.finfo(np.float64).eps)  # Base.rtoldefault(Float64)


class AbstractOperator:
    def __add__(self, other):
        return _add(self, other)

    def __radd__(self, other):
        return _add(other, self)

    def __sub__(self, other):
        return _add(self, -1 * other)

    def __rsub__(self, other):
        return _add(other, -1 * self)

    def __neg__(self):
        return -1 * self

    def __pos__(self):
        return self


class NullOperator(AbstractOperator):
    """null_operator.jl: the additive zero."""

    terms = ()

    def __mul__(self, other):
        return self

    __rmul__ = __mul__

    def __truediv__(self, other):
        return self

    def __eq__(self, other):
        return isinstance(other, NullOperator)

    def __hash__(self):
        return 0

    def __repr__(self):
        return "NullOperator()"


class PureOperator(AbstractOperator):
    """PureOperator(bitmask, bitrow, bitcol, amplitude): pure_operator.jl:25-52."""

    __slots__ = ("bitmask", "bitrow", "bitcol", "amplitude")

    def __init__(self, bitmask: int, bitrow: int, bitcol: int, amplitude: Number = 1):
        bitmask, bitrow, bitcol = int(bitmask), int(bitrow), int(bitcol)
        if (~bitmask) & bitrow:
            raise ArgumentError("every bit of bitrow not in bitmask should be set to zero")
        if (~bitmask) & bitcol:
            raise ArgumentError("every bit of bitcol not in bitmask should be set to zero")
        self.bitmask, self.bitrow, self.bitcol, self.amplitude = bitmask, bitrow, bitcol, amplitude

    @property
    def terms(self):
        return (self,)

    def _key(self):
        a = complex(self.amplitude)
        return (self.bitmask, self.bitrow, self.bitcol, a.real, a.imag)

    def __eq__(self, other):
        return isinstance(other, PureOperator) and (self.bitmask, self.bitrow, self.bitcol) == (
            other.bitmask,
            other.bitrow,
            other.bitcol,
        ) and self.amplitude == other.amplitude

    def __hash__(self):
        return hash(self._key())

    def __lt__(self, other):  # isless, pure_operator.jl:78-90
        return self._key() < other._key()

    def __mul__(self, other):
        if isinstance(other, Number):
            return PureOperator(self.bitmask, self.bitrow, self.bitcol, self.amplitude * other)
        if isinstance(other, NullOperator):
            return other
        if isinstance(other, PureOperator):
            return _pure_product(self, other)
        if isinstance(other, SumOperator):
            return SumOperator([t for t in (self * r for r in other.terms) if not isinstance(t, NullOperator)])
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, Number):
            return PureOperator(self.bitmask, self.bitrow, self.bitcol, other * self.amplitude)
        return NotImplemented

    def __truediv__(self, other):
        return PureOperator(self.bitmask, self.bitrow, self.bitcol, self.amplitude / other)

    def transpose(self):
        return PureOperator(self.bitmask, self.bitcol, self.bitrow, self.amplitude)

    def adjoint(self):
        return PureOperator(self.bitmask, self.bitcol, self.bitrow, complex(self.amplitude).conjugate())

    def __repr__(self):
        return f"PureOperator(0b{self.bitmask:b}, 0b{self.bitrow:b}, 0b{self.bitcol:b}, {self.amplitude})"


def _pure_product(lhs: PureOperator, rhs: PureOperator):
    """PureOperator * PureOperator: pure_operator.jl:136-154."""
    onlylhs = lhs.bitmask & ~rhs.bitmask
    onlyrhs = ~lhs.bitmask & rhs.bitmask
    inter = lhs.bitmask & rhs.bitmask
    union = lhs.bitmask | rhs.bitmask
    if (lhs.bitcol & inter) != (rhs.bitrow & inter):
        return NullOperator()
    return PureOperator(
        union,
        lhs.bitrow | (rhs.bitrow & onlyrhs),
        rhs.bitcol | (lhs.bitcol & onlylhs),
        lhs.amplitude * rhs.amplitude,
    )


class SumOperator(AbstractOperator):
    """SumOperator(terms): sum_operator.jl:13-23."""

    def __init__(self, terms):
        self.terms = list(terms)

    def __eq__(self, other):
        return isinstance(other, SumOperator) and self.terms == other.terms

    def __hash__(self):
        return hash(tuple(self.terms))

    def __mul__(self, other):
        if isinstance(other, Number):
            return SumOperator([t * other for t in self.terms])
        if isinstance(other, NullOperator):
            return other
        if isinstance(other, PureOperator):
            return SumOperator([x for x in (t * other for t in self.terms) if not isinstance(x, NullOperator)])
        if isinstance(other, SumOperator):
            out = []
            # vec(terms) of a matrix comprehension [tl*tr for tl in L, tr in R]
            # is column-major: the left index runs fastest (sum_operator.jl:74-78)
            for tr in other.terms:
                for tl in self.terms:
                    x = tl * tr
                    if not isinstance(x, NullOperator):
                        out.append(x)
            return SumOperator(out)
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, Number):
            return SumOperator([other * t for t in self.terms])
        return NotImplemented

    def __truediv__(self, other):
        return SumOperator([t / other for t in self.terms])

    def __repr__(self):
        return f"SumOperator({self.terms!r})"


def _as_terms(x):
    if isinstance(x, NullOperator):
        return []
    if isinstance(x, (PureOperator, SumOperator)):
        return list(x.terms)
    raise TypeError(f"not an operator: {x!r}")


def _add(lhs, rhs):
    """Operator sums (sum_operator.jl:83-123); a number c adds c * identity."""
    if isinstance(lhs, Number):
        lhs = PureOperator(0, 0, 0, lhs)
    if isinstance(rhs, Number):
        rhs = PureOperator(0, 0, 0, rhs)
    if isinstance(lhs, NullOperator):
        return rhs
    if isinstance(rhs, NullOperator):
        return lhs
    return SumOperator(_as_terms(lhs) + _as_terms(rhs))


def _is_real_number(a) -> bool:
    return not isinstance(a, complex)


def simplify(op, tol: float = RTOL_DEFAULT):
    """simplify: operator_simplify.jl:9-81.

    Sort the terms by (bitmask, bitrow, bitcol, re, im), merge equal
    (bitmask, bitrow, bitcol), drop |amplitude| <= tol, and demote to real
    amplitudes when every imaginary part is within tol of zero.
    """
    if isinstance(op, NullOperator):
        return op
    if isinstance(op, PureOperator):
        a = op.amplitude
        if abs(a) <= tol:
            return NullOperator()
        if isinstance(a, complex) and abs(a.imag) <= tol:
            return PureOperator(op.bitmask, op.bitrow, op.bitcol, a.real)
        return op
    terms = []
    for t in op.terms:
        s = simplify(t, tol)
        if not isinstance(s, NullOperator):
            # each term is simplified, then converted back to the sum's scalar
            # type: a complex amplitude whose imaginary part was dropped keeps
            # an exact zero imaginary part
            a = s.amplitude
            if isinstance(t.amplitude, complex) and not isinstance(a, complex):
                a = complex(a)
            terms.append(PureOperator(t.bitmask, t.bitrow, t.bitcol, a))
    if not terms:
        return NullOperator()
    terms.sort(key=lambda t: t._key())
    merged = []
    bm, br, bc, am = terms[0].bitmask, terms[0].bitrow, terms[0].bitcol, terms[0].amplitude
    for t in terms[1:]:
        if (bm, br, bc) == (t.bitmask, t.bitrow, t.bitcol):
            am = am + t.amplitude
        else:
            if not abs(am) <= tol:
                merged.append(PureOperator(bm, br, bc, am))
            bm, br, bc, am = t.bitmask, t.bitrow, t.bitcol, t.amplitude
    if not abs(am) <= tol:
        merged.append(PureOperator(bm, br, bc, am))
    if not merged:
        return NullOperator()
    any_complex = any(isinstance(t.amplitude, complex) for t in merged)
    if any_complex and max(abs(complex(t.amplitude).imag) for t in merged) <= tol:
        merged = [PureOperator(t.bitmask, t.bitrow, t.bitcol, complex(t.amplitude).real) for t in merged]
    if len(merged) == 1:
        return merged[0]
    return SumOperator(merged)


def pure_operator(hs, isite: int, istate_row: int, istate_col: int, amplitude: Number = 1):
    """pure_operator(hs, isite, istate_row, istate_col, amplitude): pure_operator.jl:184-201.

    isite, istate_row and istate_col are 0-based here.
    """
    hs = basespace(hs)
    site = hs.sites[isite]
    if not (0 <= istate_row < site.dimension and 0 <= istate_col < site.dimension):
        raise IndexError("state index out of range")
    bm = hs.get_bitmask(isite)
    off = hs.bitoffsets[isite]
    return PureOperator(bm, istate_row << off, istate_col << off, amplitude)


def pauli_matrix(hs, isite: int, j: str):
    """pauli_matrix(hs, isite, j): src/Toolkit/spin_half.jl:28-42 (isite 0-based).

    j in {"x", "y", "z", "+", "-"}; the local state 0 is Up, 1 is Dn.
    """
    if j == "x":
        return pure_operator(hs, isite, 0, 1, 1) + pure_operator(hs, isite, 1, 0, 1)
    if j == "y":
        return pure_operator(hs, isite, 0, 1, -1j) + pure_operator(hs, isite, 1, 0, 1j)
    if j == "z":
        return pure_operator(hs, isite, 0, 0, 1) + pure_operator(hs, isite, 1, 1, -1)
    if j == "+":
        return pure_operator(hs, isite, 0, 1, 1)
    if j == "-":
        return pure_operator(hs, isite, 1, 0, 1)
    raise ArgumentError(f"pauli matrix of type {j} not supported")


def spin_half_system(n_sites: int):
    """spin_half_system(n): src/Toolkit/spin_half.jl:9-20 -> (hs, pauli)."""
    from .hilbert import Site, State

    if n_sites > 64:
        raise ArgumentError(f"spin half system of {n_sites} sites cannot be expressed using UInt64")
    site = Site([State("Up", 1), State("Dn", -1)])
    hs = HilbertSpace([site for _ in range(n_sites)])
    return hs, (lambda i, j: pauli_matrix(hs, i, j))


def terms_arrays(op):
    """Term arrays (mask, row, col, amplitude) of an operator, in term order."""
    ts = _as_terms(op) if not isinstance(op, NullOperator) else []
    mask = np.array([t.bitmask for t in ts], dtype=np.uint64)
    row = np.array([t.bitrow for t in ts], dtype=np.uint64)
    col = np.array([t.bitcol for t in ts], dtype=np.uint64)
    amp = np.array([complex(t.amplitude) for t in ts], dtype=np.complex128)
    return mask, row, col, amp


def is_real_operator(op) -> bool:
    return all(not isinstance(t.amplitude, complex) or t.amplitude.imag == 0 for t in _as_terms(op))


def heisenberg_bond(hs, i: int, j: int, J: float = 1.0, Delta: float = 1.0):
    """J (S^x_i S^x_j + S^y_i S^y_j + Delta S^z_i S^z_j) in S = sigma/2 units,
    built from the Pauli matrices and simplified: J/2 (s+ s- + s- s+) + J Delta/4 sz sz."""
    sx = pauli_matrix(hs, i, "x") * pauli_matrix(hs, j, "x")
    sy = pauli_matrix(hs, i, "y") * pauli_matrix(hs, j, "y")
    sz = pauli_matrix(hs, i, "z") * pauli_matrix(hs, j, "z")
    return simplify(0.25 * J * (sx + sy) + 0.25 * J * Delta * sz)
