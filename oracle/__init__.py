"""CPU oracle for the symmetry-reduced Hamiltonian-application path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference algorithms of
ExactDiagonalization.jl v0.14.3 on the CPU so that the HIP product path can be
checked against them.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, always as the checker (or
the timed CPU baseline), never as the thing that is measured or shipped.

Layout
------
* ``oracle.c`` (compiled to ``liboracle.so`` by ``oracle/Makefile``): sorted
  lookup, ``symmetry_apply``, ``symmetry_reduce_serial``, row/column iterators
  of the plain and reduced representations, ``apply_serial!`` in both forms,
  ``sparse_serial`` with ``choptol!``, ``get_element`` and the vector
  reduce/unreduce maps.  Each C function cites the reference lines it follows.
* this module: the sector basis dynamic program (``hs_get_basis_list``,
  ``src/Representation/hilbert_space_representation.jl:127-206``) and
  ``quantum_number_sectors`` (``src/HilbertSpace/hilbert_space.jl:119-129``)
  in numpy, ctypes wrappers of the C part, and small dense/ARPACK eigen
  helpers used as eigenvalue references.

Pinning (which check fixes which function) is stated in the tests that use
it: ``tests/test_oracle_pins.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

RTOL_DEFAULT = float(np.sqrt(np.finfo(np.float64).eps))  # Base.rtoldefault(Float64)


def build() -> str:
    """Compile oracle.c into liboracle.so (gcc, -O2)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i64, u64p, i64p, dp, i32p = (
            ctypes.c_int64,
            ctypes.POINTER(ctypes.c_uint64),
            ctypes.POINTER(ctypes.c_int64),
            ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_int32),
        )
        L.orc_lookup.restype = i64
        L.orc_lookup.argtypes = [u64p, i64, ctypes.c_uint64]
        L.orc_lookup_many.restype = None
        L.orc_lookup_many.argtypes = [u64p, i64, i64, u64p, i64p]
        L.orc_validate_sorted_unique.restype = ctypes.c_int
        L.orc_validate_sorted_unique.argtypes = [u64p, i64]
        L.orc_symmetry_apply.restype = ctypes.c_uint64
        L.orc_symmetry_apply.argtypes = [ctypes.c_int, i32p, i32p, i32p, ctypes.c_uint64]
        L.orc_symmetry_reduce.restype = i64
        L.orc_symmetry_reduce.argtypes = [
            i64, u64p, ctypes.c_int, i32p, i32p, ctypes.c_int, i32p, dp, dp, ctypes.c_double,
            u64p, i64p, dp, dp,
        ]
        L.orc_symmetry_reduce_x.restype = i64
        L.orc_symmetry_reduce_x.argtypes = [
            i64, u64p, ctypes.c_int, i32p, i32p, ctypes.c_int, i32p, u64p, dp, dp, ctypes.c_double,
            u64p, i64p, dp, dp,
        ]
        opr_args = [i64, u64p, ctypes.c_int, i64, u64p, i64p, dp, dp, i64, u64p, u64p, u64p, dp, dp]
        L.orc_row_iterator.restype = i64
        L.orc_row_iterator.argtypes = opr_args + [i64, i64p, dp, dp]
        L.orc_col_iterator.restype = i64
        L.orc_col_iterator.argtypes = opr_args + [i64, i64p, dp, dp]
        L.orc_apply.restype = None
        L.orc_apply.argtypes = opr_args + [ctypes.c_int, dp, dp]
        L.orc_sparse.restype = i64
        L.orc_sparse.argtypes = opr_args + [ctypes.c_double, i64p, i64p, dp, dp, i64]
        pp = ctypes.POINTER
        L.orc_sparse_parallel.restype = i64
        L.orc_sparse_parallel.argtypes = opr_args + [ctypes.c_double, ctypes.c_int, pp(i64p), pp(i64p), pp(dp),
                                                     pp(dp)]
        L.orc_free.restype = None
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_get_element.restype = None
        L.orc_get_element.argtypes = opr_args + [i64, i64, dp, dp]
        L.orc_vector_reduce.restype = None
        L.orc_vector_reduce.argtypes = [i64, i64p, dp, dp, dp, dp]
        L.orc_vector_unreduce.restype = None
        L.orc_vector_unreduce.argtypes = [i64, i64p, dp, dp, dp, dp]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


# ---------------------------------------------------------------------------
# Basis (hilbert_space_representation.jl:109-206, hilbert_space.jl:119-129)
# ---------------------------------------------------------------------------

def quantum_number_sectors(site_qns):
    """quantum_number_sectors(hs): src/HilbertSpace/hilbert_space.jl:119-129.

    ``site_qns[i][j]`` is the QN tuple of state j of site i.
    """
    qdim = len(site_qns[0][0]) if site_qns else 1
    qns = {tuple([0] * qdim)}
    for states in site_qns:
        qns = {tuple(a + b for a, b in zip(q, s)) for s in states for q in qns}
    return sorted(qns)


def bit_layout(site_qns):
    """bitwidths/bitoffsets of HilbertSpace (hilbert_space.jl:30-33, site.jl:93)."""
    widths = [int(np.ceil(np.log2(len(states)))) if len(states) > 1 else 0 for states in site_qns]
    offsets = [0]
    for w in widths:
        offsets.append(offsets[-1] + w)
    return widths, offsets


def hs_get_basis_list(site_qns, allowed):
    """hs_get_basis_list(hss, BR): src/Representation/hilbert_space_representation.jl:127-206.

    Site-by-site extension of per-QN lists (``s | (i_state-1) << offset``),
    then the merge of the per-QN lists.  ``allowed`` is the sector's
    allowed-QN set (already intersected with the reachable sectors, as
    HilbertSpaceSector does at hilbert_space_sector.jl:20-63).
    """
    n = len(site_qns)
    widths, offsets = bit_layout(site_qns)
    if offsets[-1] >= 64:
        raise ValueError("UInt64 cannot represent this space")
    allowed = {tuple(q) for q in allowed}
    if not allowed & set(quantum_number_sectors(site_qns)):
        return np.zeros(0, dtype=np.uint64)
    qdim = len(site_qns[0][0])
    zero = tuple([0] * qdim)
    qn_possible = [None] * (n + 1)
    qn_possible[0] = [zero]
    for i in range(n):
        qn_possible[i + 1] = sorted(
            {tuple(a + b for a, b in zip(qa, qi)) for qi in site_qns[i] for qa in qn_possible[i]}
        )
    qn_requested = [None] * (n + 1)
    qn_requested[n] = sorted(allowed)
    for i in range(n - 1, -1, -1):
        qn_requested[i] = sorted(
            {tuple(a - b for a, b in zip(qa, qi)) for qi in site_qns[i] for qa in qn_requested[i + 1]}
        )
    schedule = [sorted(set(x) & set(y)) for x, y in zip(qn_requested, qn_possible)]
    sector = {zero: np.zeros(1, dtype=np.uint64)}
    for i in range(n):
        new = {}
        for q in schedule[i + 1]:
            parts = []
            for istate, qc in enumerate(site_qns[i]):
                qp = tuple(a - b for a, b in zip(q, qc))
                if qp in sector:
                    parts.append(sector[qp] | np.uint64(istate << offsets[i]))
            new[q] = np.concatenate(parts) if parts else np.zeros(0, dtype=np.uint64)
        sector = new
    basis = np.zeros(0, dtype=np.uint64)
    for q in list(sector.keys()):
        # merge_vec of two sorted, disjoint vectors (src/util.jl:62-86)
        basis = np.sort(np.concatenate([basis, sector[q]]), kind="mergesort")
    assert np.all(basis[1:] > basis[:-1]) if basis.size > 1 else True
    return basis


def spin_half_qns(n_sites):
    """Site([State("Up", 1), State("Dn", -1)]) repeated (spin_half.jl:9-20)."""
    return [[(1,), (-1,)] for _ in range(n_sites)]


def sector_basis(n_sites, sz_total):
    """represent(HilbertSpaceSector(hs, Sz)).basis_list for spin-1/2 sites."""
    return hs_get_basis_list(spin_half_qns(n_sites), {(sz_total,)})


# ---------------------------------------------------------------------------
# Wrappers of the C restatement
# ---------------------------------------------------------------------------

def lookup(keys, queries):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    q = np.ascontiguousarray(queries, dtype=np.uint64)
    out = np.empty(q.size, dtype=np.int64)
    lib().orc_lookup_many(_p(keys, ctypes.c_uint64), keys.size, q.size, _p(q, ctypes.c_uint64), _p(out, ctypes.c_int64))
    return out


def validate_sorted_unique(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    return int(lib().orc_validate_sorted_unique(_p(keys, ctypes.c_uint64), keys.size))


def symmetry_apply(n_sites, perm_map, b, bitwidths=None):
    """perm_map: 0-based image site of each site (SitePermutation.map - 1)."""
    widths = np.ones(n_sites, dtype=np.int32) if bitwidths is None else np.asarray(bitwidths, np.int32)
    offsets = np.concatenate([[0], np.cumsum(widths)]).astype(np.int32)
    m = np.ascontiguousarray(perm_map, dtype=np.int32)
    return int(
        lib().orc_symmetry_apply(n_sites, _p(offsets, ctypes.c_int32), _p(widths, ctypes.c_int32), _p(m, ctypes.c_int32), int(b))
    )


class Reduced:
    """Output of symmetry_reduce_serial (reduced_hilbert_space_representation.jl:13-22)."""

    def __init__(self, parent, basis_list, mapping_index, mapping_amplitude):
        self.parent = parent
        self.basis_list = basis_list
        self.basis_mapping_index = mapping_index
        self.basis_mapping_amplitude = mapping_amplitude

    @property
    def dimension(self):
        return int(self.basis_list.size)


def symmetry_reduce(parent, n_sites, perms, chars, tol=RTOL_DEFAULT, bitwidths=None, flips=None):
    """symmetry_reduce_serial(hsr, symops_and_amplitudes; tol) (symmetry_reduce_generic.jl:22-106).

    perms: (G, n_sites) 0-based image sites, identity first; chars: complex (G,).
    flips: optional (G,) words XORed into each permuted word (GlobalBitFlip
    composed with the permutation, bitflipsymmetry.jl:26-38).
    """
    parent = np.ascontiguousarray(parent, dtype=np.uint64)
    perms = np.ascontiguousarray(perms, dtype=np.int32)
    G = perms.shape[0]
    chars = np.asarray(chars, dtype=np.complex128)
    cre = np.ascontiguousarray(chars.real)
    cim = np.ascontiguousarray(chars.imag)
    widths = np.ones(n_sites, dtype=np.int32) if bitwidths is None else np.asarray(bitwidths, np.int32)
    offsets = np.concatenate([[0], np.cumsum(widths)]).astype(np.int32)
    n = parent.size
    red = np.empty(max(n, 1), dtype=np.uint64)
    midx = np.empty(max(n, 1), dtype=np.int64)
    are = np.empty(max(n, 1), dtype=np.float64)
    aim = np.empty(max(n, 1), dtype=np.float64)
    fl = None if flips is None else np.ascontiguousarray(np.asarray(flips, dtype=np.uint64))
    d = lib().orc_symmetry_reduce_x(
        n, _p(parent, ctypes.c_uint64), n_sites, _p(offsets, ctypes.c_int32), _p(widths, ctypes.c_int32), G,
        _p(perms, ctypes.c_int32), None if fl is None else _p(fl, ctypes.c_uint64), _p(cre, ctypes.c_double),
        _p(cim, ctypes.c_double), float(tol), _p(red, ctypes.c_uint64), _p(midx, ctypes.c_int64),
        _p(are, ctypes.c_double), _p(aim, ctypes.c_double),
    )
    if d == -1:
        raise ValueError("all amplitudes need to have norm 1")
    if d == -2:
        raise KeyError("symmetry image outside the parent basis")
    return Reduced(parent, red[:d].copy(), midx[:n].copy(), (are[:n] + 1j * aim[:n]))


class Opr:
    """An operator representation for the C oracle (plain or reduced)."""

    def __init__(self, space, terms):
        """space: parent basis (np.uint64 array) or a Reduced; terms: (mask,row,col,amp) arrays."""
        mask, row, col, amp = terms
        self.mask = np.ascontiguousarray(mask, dtype=np.uint64)
        self.row = np.ascontiguousarray(row, dtype=np.uint64)
        self.col = np.ascontiguousarray(col, dtype=np.uint64)
        amp = np.asarray(amp, dtype=np.complex128)
        self.are = np.ascontiguousarray(amp.real)
        self.aim = np.ascontiguousarray(amp.imag)
        if isinstance(space, Reduced):
            self.reduced = 1
            self.parent = np.ascontiguousarray(space.parent, dtype=np.uint64)
            self.red = np.ascontiguousarray(space.basis_list, dtype=np.uint64)
            self.midx = np.ascontiguousarray(space.basis_mapping_index, dtype=np.int64)
            self.mre = np.ascontiguousarray(space.basis_mapping_amplitude.real)
            self.mim = np.ascontiguousarray(space.basis_mapping_amplitude.imag)
        else:
            self.reduced = 0
            self.parent = np.ascontiguousarray(space, dtype=np.uint64)
            self.red = np.zeros(1, dtype=np.uint64)
            self.midx = np.zeros(1, dtype=np.int64)
            self.mre = np.zeros(1)
            self.mim = np.zeros(1)
        self.dim = int(self.red.size if self.reduced else self.parent.size)

    def _args(self):
        return (
            self.parent.size, _p(self.parent, ctypes.c_uint64), self.reduced,
            self.red.size if self.reduced else 0, _p(self.red, ctypes.c_uint64),
            _p(self.midx, ctypes.c_int64), _p(self.mre, ctypes.c_double), _p(self.mim, ctypes.c_double),
            self.mask.size, _p(self.mask, ctypes.c_uint64), _p(self.row, ctypes.c_uint64),
            _p(self.col, ctypes.c_uint64), _p(self.are, ctypes.c_double), _p(self.aim, ctypes.c_double),
        )

    def row_iterator(self, irow):
        T = max(self.mask.size, 1)
        idx = np.empty(T, np.int64)
        vr = np.empty(T)
        vi = np.empty(T)
        n = lib().orc_row_iterator(*self._args(), irow, _p(idx, ctypes.c_int64), _p(vr, ctypes.c_double), _p(vi, ctypes.c_double))
        return idx[:n].copy(), vr[:n] + 1j * vi[:n]

    def col_iterator(self, icol):
        T = max(self.mask.size, 1)
        idx = np.empty(T, np.int64)
        vr = np.empty(T)
        vi = np.empty(T)
        n = lib().orc_col_iterator(*self._args(), icol, _p(idx, ctypes.c_int64), _p(vr, ctypes.c_double), _p(vi, ctypes.c_double))
        return idx[:n].copy(), vr[:n] + 1j * vi[:n]

    def apply(self, out, state, transpose=False):
        """out += A @ state (transpose=False) or out += state @ A; complex vectors."""
        x = np.ascontiguousarray(state, dtype=np.complex128)
        y = np.ascontiguousarray(out, dtype=np.complex128)
        if x.size != self.dim or y.size != self.dim:
            raise ValueError("DimensionMismatch")
        lib().orc_apply(*self._args(), 1 if transpose else 0, _p(x.view(np.float64), ctypes.c_double), _p(y.view(np.float64), ctypes.c_double))
        return y

    def matvec(self, state):
        return self.apply(np.zeros(self.dim, np.complex128), state)

    def sparse(self, tol=RTOL_DEFAULT):
        """(colptr, rowval, nzval) 0-based CSC, as sparse_serial."""
        cap = max(self.dim * max(self.mask.size, 1), 1)
        colptr = np.empty(self.dim + 1, np.int64)
        rowval = np.empty(cap, np.int64)
        nr = np.empty(cap)
        ni = np.empty(cap)
        nnz = lib().orc_sparse(*self._args(), float(tol), _p(colptr, ctypes.c_int64), _p(rowval, ctypes.c_int64), _p(nr, ctypes.c_double), _p(ni, ctypes.c_double), cap)
        assert nnz >= 0
        return colptr, rowval[:nnz].copy(), nr[:nnz] + 1j * ni[:nnz]

    def sparse_parallel(self, tol=RTOL_DEFAULT, n_threads=None):
        """(colptr, rowval, nzval) 0-based CSC, as sparse_parallel (column
        blocks over n_threads OpenMP threads, default os.cpu_count())."""
        nt = int(n_threads or os.cpu_count() or 1)
        i64p = ctypes.POINTER(ctypes.c_int64)
        dp = ctypes.POINTER(ctypes.c_double)
        cp, rv, nr, ni = i64p(), i64p(), dp(), dp()
        nnz = lib().orc_sparse_parallel(*self._args(), float(tol), nt, ctypes.byref(cp), ctypes.byref(rv),
                                        ctypes.byref(nr), ctypes.byref(ni))
        if nnz < 0:
            raise MemoryError("orc_sparse_parallel: allocation failed")
        try:
            colptr = np.ctypeslib.as_array(cp, shape=(self.dim + 1,)).copy()
            rowval = np.ctypeslib.as_array(rv, shape=(max(nnz, 1),))[:nnz].copy()
            nz = np.ctypeslib.as_array(nr, shape=(max(nnz, 1),))[:nnz] + 1j * np.ctypeslib.as_array(ni, shape=(max(nnz, 1),))[:nnz]
        finally:
            for q in (cp, rv, nr, ni):
                lib().orc_free(ctypes.cast(q, ctypes.c_void_p))
        return colptr, rowval, nz

    def scipy_sparse(self, tol=RTOL_DEFAULT):
        import scipy.sparse as sp

        colptr, rowval, nz = self.sparse(tol)
        return sp.csc_matrix((nz, rowval, colptr), shape=(self.dim, self.dim))

    def get_element(self, i, j):
        re = ctypes.c_double()
        im = ctypes.c_double()
        lib().orc_get_element(*self._args(), i, j, ctypes.byref(re), ctypes.byref(im))
        return complex(re.value, im.value)

    def dense(self):
        """Matrix(opr): abstract_operator_representation.jl:121-132 (column iterators)."""
        m = np.zeros((self.dim, self.dim), np.complex128)
        for c in range(self.dim):
            idx, v = self.col_iterator(c)
            for k in range(idx.size):
                if 0 <= idx[k] < self.dim:
                    m[idx[k], c] += v[k]
        return m


def vector_reduce(red: Reduced, large):
    large = np.ascontiguousarray(large, dtype=np.complex128)
    small = np.zeros(red.dimension, np.complex128)
    midx = np.ascontiguousarray(red.basis_mapping_index, np.int64)
    mre = np.ascontiguousarray(red.basis_mapping_amplitude.real)
    mim = np.ascontiguousarray(red.basis_mapping_amplitude.imag)
    lib().orc_vector_reduce(midx.size, _p(midx, ctypes.c_int64), _p(mre, ctypes.c_double), _p(mim, ctypes.c_double), _p(large.view(np.float64), ctypes.c_double), _p(small.view(np.float64), ctypes.c_double))
    return small


def vector_unreduce(red: Reduced, small):
    small = np.ascontiguousarray(small, dtype=np.complex128)
    large = np.zeros(red.parent.size, np.complex128)
    midx = np.ascontiguousarray(red.basis_mapping_index, np.int64)
    mre = np.ascontiguousarray(red.basis_mapping_amplitude.real)
    mim = np.ascontiguousarray(red.basis_mapping_amplitude.imag)
    lib().orc_vector_unreduce(midx.size, _p(midx, ctypes.c_int64), _p(mre, ctypes.c_double), _p(mim, ctypes.c_double), _p(small.view(np.float64), ctypes.c_double), _p(large.view(np.float64), ctypes.c_double))
    return large


# ---------------------------------------------------------------------------
# Eigenvalue references (the reference's callers use Arpack.eigs, which wraps
# ARPACK's dsaupd/znaupd; scipy.sparse.linalg.eigsh wraps the same ARPACK).
# ---------------------------------------------------------------------------

def lowest_eigenvalues(opr: Opr, nev=1, tol=0.0):
    import scipy.sparse.linalg as sla

    if opr.dim <= 400:
        w = np.linalg.eigvalsh(opr.dense())
        return np.sort(w)[:nev]
    A = opr.scipy_sparse()
    if np.all(A.data.imag == 0):
        A = A.real
    v0 = np.ones(opr.dim) / np.sqrt(opr.dim)
    w = sla.eigsh(A, k=nev, which="SA", tol=tol, v0=v0.astype(A.dtype), return_eigenvectors=False)
    return np.sort(w)
