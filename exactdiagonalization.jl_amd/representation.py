"""Device-backed representations: the drop-in for the reference's
HilbertSpaceRepresentation / ReducedHilbertSpaceRepresentation and
OperatorRepresentation / ReducedOperatorRepresentation.

Reference interfaces mirrored (ExactDiagonalization.jl v0.14.3):
  represent(hss)                    hilbert_space_representation.jl:215-230
  represent(hs, basis_list)         :239-256
  symmetry_reduce(hsr, symops; tol) symmetry_reduce_generic.jl:7-14
  represent(hsr | rhsr, op)         operator_representation.jl:34-36,
                                    reduced_operator_representation.jl:40-42
  apply!/apply_serial!/apply_parallel!, mul!, sparse, Matrix, get_row,
  get_column, get_element, get_row_iterator, get_column_iterator
                                    abstract_operator_representation.jl:110-409
Julia's in-place names end in "!"; here they end in "_" (apply_, mul_).
Indices are 0-based; -1 marks an invalid image as in the reference.

Vectors may be numpy arrays (copied to and from the device) or torch tensors
on a GPU (used in place).  Every computation runs in libedhip's HIP kernels.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import ArgumentError, BoundsError, DimensionMismatch, ptr
from .hilbert import HilbertSpace, HilbertSpaceSector, basespace, popcount_sets
from .operators import NullOperator, RTOL_DEFAULT, terms_arrays

# ---------------------------------------------------------------------------
# spaces
# ---------------------------------------------------------------------------


class AbstractHilbertSpaceRepresentation:
    _handle = None

    @property
    def dimension(self) -> int:
        d = ctypes.c_int64()
        self.ctx.check(_lib.lib().edhip_space_dim(self._handle, ctypes.byref(d)))
        return int(d.value)

    def __len__(self):
        return self.dimension

    @property
    def basis_list(self) -> np.ndarray:
        if self._basis is None:
            n = self.dimension
            out = np.empty(max(n, 1), dtype=np.uint64)
            self.ctx.check(_lib.lib().edhip_space_copy_basis(self._handle, ptr(out, ctypes.c_uint64)))
            self._basis = out[:n]
        return self._basis

    def lookup(self, keys) -> np.ndarray:
        """get(basis_lookup, key, -1) for each key (frozensortedarray.jl:45-48)."""
        k = np.ascontiguousarray(np.atleast_1d(keys), dtype=np.uint64)
        out = np.empty(k.size, dtype=np.int64)
        self.ctx.check(_lib.lib().edhip_space_lookup(self._handle, k.size, ptr(k, ctypes.c_uint64), ptr(out, ctypes.c_int64)))
        return out

    def bitwidth(self) -> int:
        return self.hilbert_space.bitwidth()

    def __del__(self):
        try:
            if self._handle is not None and _lib._lib is not None:
                _lib._lib.edhip_space_destroy(self._handle)
        except Exception:
            pass


class HilbertSpaceRepresentation(AbstractHilbertSpaceRepresentation):
    """hilbert_space_representation.jl:16-74 (UInt64 words, array lookup)."""

    def __init__(self, hilbert_space, handle, ctx):
        self.hilbert_space = basespace(hilbert_space)
        self._handle = handle
        self.ctx = ctx
        self._basis = None

    def __eq__(self, other):
        return (
            isinstance(other, HilbertSpaceRepresentation)
            and self.hilbert_space == other.hilbert_space
            and np.array_equal(self.basis_list, other.basis_list)
        )

    def __repr__(self):
        return f"HilbertSpaceRepresentation(dim={self.dimension})"


_BR_BITS = {"UInt8": 8, "UInt16": 16, "UInt32": 32, "UInt64": 64, "UInt": 64, np.uint8: 8, np.uint16: 16, np.uint32: 32, np.uint64: 64}


def _check_width(hs, BR):
    bits = _BR_BITS.get(BR, 64) if BR is not None else 64
    need = basespace(hs).bitwidth()
    if bits <= need:
        raise ArgumentError(f"type {BR} not enough to represent the hilbert space (need {need} bits)")


def represent(space, arg=None, *, ctx=None):
    """represent(hss[, BR]) / represent(hs, basis_list) / represent(hsr, op).

    Dispatches like the reference: a Hilbert space (or sector) gives a
    HilbertSpaceRepresentation; a representation plus an operator gives an
    (Reduced)OperatorRepresentation.
    """
    if isinstance(space, AbstractHilbertSpaceRepresentation):
        if isinstance(space, ReducedHilbertSpaceRepresentation):
            return ReducedOperatorRepresentation(space, arg)
        return OperatorRepresentation(space, arg)
    if isinstance(arg, (list, tuple, np.ndarray)):
        return represent_array(space, arg, ctx=ctx)
    return represent_array(space, None, BR=arg, ctx=ctx)


def represent_array(space, basis_list=None, *, BR=None, ctx=None):
    """represent_array(hs | hss, [basis_list]): hilbert_space_representation.jl:226-256."""
    ctx = ctx or _lib.default_context()
    _check_width(space, BR)
    nb = basespace(space).bitwidth()
    h = ctypes.c_void_p()
    if basis_list is not None:
        b = np.sort(np.asarray(basis_list, dtype=np.uint64))
        b = np.ascontiguousarray(b)
        ctx.check(_lib.lib().edhip_space_from_list(ctx.handle, nb, b.size, ptr(b, ctypes.c_uint64), ctypes.byref(h)))
    else:
        pops = popcount_sets(space)
        if pops is None:
            _sector_qn(space, ctx, h)
        else:
            p = np.ascontiguousarray(pops, dtype=np.int32)
            ctx.check(_lib.lib().edhip_space_sector(ctx.handle, nb, p.size, ptr(p, ctypes.c_int32), ctypes.byref(h)))
    return HilbertSpaceRepresentation(space, h, ctx)


def _sector_qn(space, ctx, h):
    """Device enumeration of a sector of arbitrary sites and quantum-number
    tuples (hs_get_basis_list, hilbert_space_representation.jl:109-206)."""
    base = basespace(space)
    qd = base.qntype_len()
    allowed = (space.allowed_quantum_numbers if isinstance(space, HilbertSpaceSector)
               else set(base.quantum_number_sectors()))
    offs = np.ascontiguousarray(base.bitoffsets, dtype=np.int32)
    nst = np.ascontiguousarray([len(s.states) for s in base.sites], dtype=np.int32)
    qns = np.ascontiguousarray([c for s in base.sites for st in s.states for c in st.quantum_number], dtype=np.int64)
    al = np.ascontiguousarray([c for q in sorted(allowed) for c in q], dtype=np.int64)
    for a in (qns, al):
        if a.size and (a.min() < -(2**31) or a.max() >= 2**31):
            raise ArgumentError("quantum numbers must fit in 32 bits")
    qns, al = qns.astype(np.int32), al.astype(np.int32)
    if not qns.size:
        qns = np.zeros(1, dtype=np.int32)
    if not al.size:
        al = np.zeros(1, dtype=np.int32)
    if not nst.size:
        nst = np.zeros(1, dtype=np.int32)
    ctx.check(_lib.lib().edhip_space_sector_qn(
        ctx.handle, base.n_sites, ptr(offs, ctypes.c_int32), ptr(nst, ctypes.c_int32), qd,
        ptr(qns, ctypes.c_int32), len(allowed), ptr(al, ctypes.c_int32), ctypes.byref(h)))


class SitePermutation:
    """A site permutation; ``map`` is 1-based as LatticeTools' SitePermutation:
    site i is sent to site map[i] (symmetry_apply.jl:82-92)."""

    def __init__(self, map1):
        self.map = [int(x) for x in map1]
        if sorted(self.map) != list(range(1, len(self.map) + 1)):
            raise ArgumentError("not a permutation")

    def __len__(self):
        return len(self.map)

    def __mul__(self, other):  # (self * other)(i) = self(other(i))
        if isinstance(other, (GlobalBitFlip, DirectProductOperation)):
            return DirectProductOperation(self, other)
        return SitePermutation([self.map[other.map[i] - 1] for i in range(len(self.map))])

    def __eq__(self, other):
        return isinstance(other, SitePermutation) and self.map == other.map

    def __hash__(self):
        return hash(tuple(self.map))

    def __repr__(self):
        return f"SitePermutation({self.map})"


class GlobalBitFlip:
    """GlobalBitFlip(value=False): flips every bit of the word when value is
    true (bitflipsymmetry.jl:8-38): the spin flip of a spin-1/2 system."""

    def __init__(self, value: bool = False):
        self.value = bool(value)

    def __mul__(self, other):
        if isinstance(other, GlobalBitFlip):
            return GlobalBitFlip(self.value ^ other.value)
        return DirectProductOperation(self, other)

    def __eq__(self, other):
        return isinstance(other, GlobalBitFlip) and self.value == other.value

    def __hash__(self):
        return hash(("GlobalBitFlip", self.value))

    def isidentity(self) -> bool:
        return not self.value

    def __repr__(self):
        return f"GlobalBitFlip({self.value})"


class DirectProductOperation:
    """A site permutation combined with a global bit flip (the two commute);
    symmetry_apply.jl:65-76 applies the factors right to left.  Kept in the
    normal form (permutation, flip)."""

    def __init__(self, *ops):
        perm, flip = None, False
        for op in ops:
            if isinstance(op, DirectProductOperation):
                p, f = op.perm, op.flip
            elif isinstance(op, GlobalBitFlip):
                p, f = None, op.value
            elif isinstance(op, SitePermutation):
                p, f = op, False
            else:
                raise ArgumentError(f"unsupported symmetry operation {op!r}")
            if p is not None:
                perm = p if perm is None else perm * p
            flip ^= f
        self.perm, self.flip = perm, flip

    def __mul__(self, other):
        return DirectProductOperation(self, other)

    def __eq__(self, other):
        return isinstance(other, DirectProductOperation) and self.perm == other.perm and self.flip == other.flip

    def __hash__(self):
        return hash((self.perm, self.flip))

    def __repr__(self):
        return f"DirectProductOperation({self.perm!r}, GlobalBitFlip({self.flip}))"


def _perm_and_flip(hs, op):
    """(bit permutation, xor word) of a symmetry operation."""
    hs = basespace(hs)
    nb = hs.bitwidth()
    ident = SitePermutation(list(range(1, len(hs.sites) + 1)))
    if isinstance(op, SitePermutation):
        return bit_permutation(hs, op), 0
    if isinstance(op, GlobalBitFlip):
        return bit_permutation(hs, ident), ((1 << nb) - 1) if op.value else 0
    if isinstance(op, DirectProductOperation):
        perm = op.perm if op.perm is not None else ident
        return bit_permutation(hs, perm), ((1 << nb) - 1) if op.flip else 0
    raise ArgumentError(f"unsupported symmetry operation {op!r}")


def make_product_irrep(*irreps):
    """Direct product of one-dimensional irreps given as lists of
    (operation, amplitude): every combination, operations multiplied and
    amplitudes multiplied, the identity first when each factor has it first
    (the role of LatticeTools' make_product_irrep in
    examples/spinhalf_triangular.jl:87-96)."""
    out = [(None, 1.0)]
    for irrep in irreps:
        items = list(irrep.symops_and_amplitudes()) if hasattr(irrep, "symops_and_amplitudes") else list(irrep)
        out = [(op if acc is None else acc * op, a * b) for acc, a in out for op, b in items]
    return out


def symmetry_apply(hs, perm, x):
    """symmetry_apply(hs, ::SitePermutation, bitrep | PureOperator | SumOperator):
    symmetry_apply.jl:54-106; GlobalBitFlip and DirectProductOperation on a
    word (bitflipsymmetry.jl:26-38, symmetry_apply.jl:65-76).  Host-side (used
    to validate groups and operators)."""
    from .operators import PureOperator, SumOperator

    hs = basespace(hs)
    if isinstance(perm, (GlobalBitFlip, DirectProductOperation)):
        if isinstance(x, (NullOperator, PureOperator, SumOperator)):
            raise ArgumentError("a bit flip acts on words only")
        b = int(x)
        nb = hs.bitwidth()
        if b >> nb:
            raise ArgumentError("binary representation too small")
        if isinstance(perm, DirectProductOperation):
            if perm.flip:
                b = ((1 << nb) - 1) & ~b
            return symmetry_apply(hs, perm.perm, b) if perm.perm is not None else (b, 1)
        return ((((1 << nb) - 1) & ~b) if perm.value else b, 1)
    if isinstance(x, NullOperator):
        return x
    if isinstance(x, PureOperator):
        return PureOperator(
            symmetry_apply(hs, perm, x.bitmask)[0],
            symmetry_apply(hs, perm, x.bitrow)[0],
            symmetry_apply(hs, perm, x.bitcol)[0],
            x.amplitude,
        )
    if isinstance(x, SumOperator):
        return SumOperator([symmetry_apply(hs, perm, t) for t in x.terms])
    b = int(x)
    out = 0
    for i, j in enumerate(perm.map):
        w = hs.bitwidths[i]
        out |= ((b >> hs.bitoffsets[i]) & ((1 << w) - 1)) << hs.bitoffsets[j - 1]
    return (out, 1)


def isinvariant(hs, perm: SitePermutation, op) -> bool:
    """isinvariant(hs, symop, op): symmetry_apply.jl:110-116."""
    from .operators import simplify

    return isinstance(simplify(op - symmetry_apply(hs, perm, op)), NullOperator)


def bit_permutation(hs, perm: SitePermutation) -> np.ndarray:
    """The bit-level image of each bit of the word under a site permutation."""
    hs = basespace(hs)
    nb = hs.bitwidth()
    out = np.empty(nb, dtype=np.int32)
    for i, j in enumerate(perm.map):
        if hs.bitwidths[i] != hs.bitwidths[j - 1]:
            raise ArgumentError("a site permutation must map sites of equal bit width")
        for k in range(hs.bitwidths[i]):
            out[hs.bitoffsets[i] + k] = hs.bitoffsets[j - 1] + k
    return out


class ReducedHilbertSpaceRepresentation(AbstractHilbertSpaceRepresentation):
    """reduced_hilbert_space_representation.jl:13-38 on the device.

    Holds the ascending representatives and their orbit sizes; the parent-size
    arrays basis_mapping_index / basis_mapping_amplitude of the reference are
    produced on demand (edhip_space_export_mapping) for checks only.
    """

    def __init__(self, parent: HilbertSpaceRepresentation, handle, ctx, symops_and_amplitudes):
        self.parent = parent
        self.hilbert_space = parent.hilbert_space
        self._handle = handle
        self.ctx = ctx
        self._basis = None
        self._mapping = None
        self.symops_and_amplitudes = symops_and_amplitudes

    def _export(self):
        if self._mapping is None:
            n = self.parent.dimension
            idx = np.empty(max(n, 1), dtype=np.int64)
            re = np.empty(max(n, 1))
            im = np.empty(max(n, 1))
            self.ctx.check(
                _lib.lib().edhip_space_export_mapping(
                    self._handle, ptr(idx, ctypes.c_int64), ptr(re, ctypes.c_double), ptr(im, ctypes.c_double)
                )
            )
            self._mapping = (idx[:n], re[:n] + 1j * im[:n])
        return self._mapping

    @property
    def basis_mapping_index(self) -> np.ndarray:
        return self._export()[0]

    @property
    def basis_mapping_amplitude(self) -> np.ndarray:
        return self._export()[1]

    @property
    def orbit_sizes(self) -> np.ndarray:
        n = self.dimension
        out = np.empty(max(n, 1), dtype=np.uint16)
        self.ctx.check(_lib.lib().edhip_space_copy_orbit(self._handle, ptr(out, ctypes.c_uint16)))
        return out[:n]

    def canonicaliser(self) -> str:
        r, g, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.ctx.check(_lib.lib().edhip_space_info(self._handle, ctypes.byref(r), ctypes.byref(g), ctypes.byref(k)))
        return {0: "none", 1: "generic", 2: "torus"}[k.value]

    def __repr__(self):
        return f"ReducedHilbertSpaceRepresentation(dim={self.dimension}, parent_dim={self.parent.dimension})"


def _symops(symops_and_amplitudes):
    if hasattr(symops_and_amplitudes, "symops_and_amplitudes"):
        return list(symops_and_amplitudes.symops_and_amplitudes())
    return list(symops_and_amplitudes)


def symmetry_reduce(hsr, symops_and_amplitudes, tol: float = RTOL_DEFAULT):
    """symmetry_reduce(hsr, symops_and_amplitudes; tol): symmetry_reduce_generic.jl:7-14.

    ``symops_and_amplitudes``: iterable of (SitePermutation, amplitude) with
    the identity first, or an irrep object with ``symops_and_amplitudes()``
    (the IrrepComponent route of symmetry_reduce_translation.jl, which stores
    the same conj(amplitude)/sqrt(orbit) phases).

    With a ReducedHilbertSpaceRepresentation and a vector instead, this is the
    vector map of symmetry_reduce.jl:29-84 (see symmetry_reduce_vector).
    """
    if isinstance(hsr, ReducedHilbertSpaceRepresentation):
        return symmetry_reduce_vector(hsr, symops_and_amplitudes)
    if not isinstance(hsr, HilbertSpaceRepresentation):
        raise ArgumentError("symmetry_reduce needs a HilbertSpaceRepresentation")
    ops = _symops(symops_and_amplitudes)
    if not ops:
        raise ArgumentError("empty symmetry group")
    pf = [_perm_and_flip(hsr.hilbert_space, p) for p, _ in ops]
    perms = np.ascontiguousarray(np.stack([p for p, _ in pf]), dtype=np.int32)
    flips = np.ascontiguousarray(np.array([f for _, f in pf], dtype=np.uint64))
    chi = np.array([complex(a) for _, a in ops], dtype=np.complex128)
    cre, cim = np.ascontiguousarray(chi.real), np.ascontiguousarray(chi.imag)
    h = ctypes.c_void_p()
    ctx = hsr.ctx
    ctx.check(
        _lib.lib().edhip_space_reduce_ex(
            ctx.handle, hsr._handle, len(ops), ptr(perms, ctypes.c_int32),
            ptr(flips, ctypes.c_uint64) if flips.any() else None, ptr(cre, ctypes.c_double),
            ptr(cim, ctypes.c_double), float(tol), ctypes.byref(h),
        )
    )
    return ReducedHilbertSpaceRepresentation(hsr, h, ctx, ops)


symmetry_reduce_serial = symmetry_reduce
symmetry_reduce_parallel = symmetry_reduce


def _vector_map(rhsr: ReducedHilbertSpaceRepresentation, vec, reduce: bool):
    """Device vector map between rhsr.parent and rhsr.  The result type is
    promote_type(C, eltype(vec)) widened to 64-bit components, with C complex
    when the amplitudes are (symmetry_reduce.jl:43,215).  GPU tensors give a
    GPU tensor, anything else a numpy array."""
    import torch

    if not isinstance(rhsr, ReducedHilbertSpaceRepresentation):
        raise ArgumentError("expected a ReducedHilbertSpaceRepresentation")
    amps = np.array([a for _, a in _symops(rhsr.symops_and_amplitudes)])
    on_device = _is_torch(vec)
    if on_device:
        if not vec.is_cuda:
            raise ArgumentError("a tensor argument must live on the GPU")
        cx = np.iscomplexobj(amps) or vec.is_complex()
        src = vec.reshape(-1)
    else:
        arr = np.asarray(vec)
        cx = np.iscomplexobj(amps) or np.iscomplexobj(arr)
        src = torch.from_numpy(np.ascontiguousarray(arr.reshape(-1)))
    dt = torch.complex128 if cx else torch.float64
    x = src.to(device=f"cuda:{rhsr.ctx.device}", dtype=dt).contiguous()
    n_out = rhsr.dimension if reduce else rhsr.parent.dimension
    y = torch.empty(max(n_out, 1), dtype=dt, device=x.device)
    prec = _lib.C128 if cx else _lib.F64
    fn = _lib.lib().edhip_space_vector_reduce if reduce else _lib.lib().edhip_space_vector_unreduce
    _device_call(rhsr.ctx, fn, rhsr._handle, prec, x.numel(), ctypes.c_void_p(x.data_ptr()), n_out,
                 ctypes.c_void_p(y.data_ptr()))
    y = y[:n_out]
    return y if on_device else y.cpu().numpy()


def symmetry_reduce_vector(rhsr: ReducedHilbertSpaceRepresentation, large_vector):
    """symmetry_reduce(rhsr, large_vector) (symmetry_reduce.jl:29-56):
    small[r] = sum_p conj(U[p]) large[p] over the parent words p of the orbit
    of r.  DimensionMismatch unless len(large_vector) == dimension(parent)."""
    return _vector_map(rhsr, large_vector, True)


def symmetry_unreduce(rhsr: ReducedHilbertSpaceRepresentation, small_vector):
    """symmetry_unreduce(rhsr, small_vector) (symmetry_reduce.jl:208-225):
    large[p] = U[p] small[r(p)], 0 outside the reduced space.
    DimensionMismatch unless len(small_vector) == dimension(rhsr)."""
    return _vector_map(rhsr, small_vector, False)


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _device_call(ctx, fn, *args):
    """Run a library call on GPU tensors in stream order, without host
    synchronisation: the call is queued after the caller's current torch
    stream (the producers of its inputs), and that stream waits for it (its
    outputs are complete for whatever the caller queues next).  The memory of
    the inputs stays valid meanwhile: later work on the caller's stream,
    which may reuse it once it is freed, is queued after the call."""
    import torch

    cur = torch.cuda.current_stream(ctx.device)
    run = ctx.order_after(cur)
    ctx.check(fn(*args))
    if run is not cur:
        cur.wait_stream(run)


def _torch_prec(t):
    import torch

    return {torch.float64: _lib.F64, torch.complex128: _lib.C128, torch.float32: _lib.F32, torch.complex64: _lib.C64}.get(t.dtype)


class AbstractOperatorRepresentation:
    """AbstractOperatorRepresentation{S} <: AbstractMatrix{S}
    (abstract_operator_representation.jl:16-46)."""

    # numpy defers `vector @ opr` to __rmatmul__
    __array_ufunc__ = None

    def __init__(self, space, operator):
        if operator is None:
            operator = NullOperator()
        self.space = space
        self.operator = operator
        self.ctx = space.ctx
        mask, row, col, amp = terms_arrays(operator)
        self._n_terms = mask.size
        are = np.ascontiguousarray(amp.real)
        aim = np.ascontiguousarray(amp.imag)
        h = ctypes.c_void_p()
        self.ctx.check(
            _lib.lib().edhip_op_create(
                self.ctx.handle, space._handle, mask.size, ptr(mask, ctypes.c_uint64), ptr(row, ctypes.c_uint64),
                ptr(col, ctypes.c_uint64), ptr(are, ctypes.c_double), ptr(aim, ctypes.c_double), ctypes.byref(h),
            )
        )
        self._handle = h
        self._sparse_cache = {}

    def __del__(self):
        try:
            if self._handle is not None and _lib._lib is not None:
                _lib._lib.edhip_op_destroy(self._handle)
        except Exception:
            pass

    # ---- traits ------------------------------------------------------------
    def get_space(self):
        return self.space

    @property
    def shape(self):
        d = self.space.dimension
        return (d, d)

    def size(self, dim=None):
        if dim is None:
            return self.shape
        if dim not in (1, 2):
            raise BoundsError("size")
        return self.shape[dim - 1]

    @property
    def dimension(self):
        return self.space.dimension

    def bitwidth(self):
        return self.space.bitwidth()

    @property
    def is_complex(self) -> bool:
        c = ctypes.c_int()
        self.ctx.check(_lib.lib().edhip_op_is_complex(self._handle, ctypes.byref(c)))
        return bool(c.value)

    @property
    def dtype(self):
        return np.dtype(np.complex128) if self.is_complex else np.dtype(np.float64)

    # ---- strategy (performance knob; every choice gives the same matrix) -------
    _KERNELS = {"auto": 0, "terms": 1, "pairs": 2, "assembled": 3}
    _LOOKUPS = {"auto": 0, "search": 1, "hash": 2, "bitmap": 3, "xspread": 4}

    def set_strategy(self, kernel: str = "auto", lookup: str = "auto"):
        self.ctx.check(_lib.lib().edhip_op_set_strategy(self._handle, self._KERNELS[kernel], self._LOOKUPS[lookup]))

    def last_strategy(self):
        k, l = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(_lib.lib().edhip_op_last_strategy(self._handle, ctypes.byref(k), ctypes.byref(l)))
        return ({0: None, 1: "terms", 2: "pairs", 3: "assembled"}[k.value],
                {0: None, 1: "search", 2: "hash", 3: "bitmap", 4: "xspread"}[l.value])

    def assembled_info(self):
        """(entries, device bytes, build seconds) of the assembled rows the
        operator holds (kernel "assembled"); zeros when it holds none."""
        n, b, t = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        self.ctx.check(_lib.lib().edhip_op_assembled_info(self._handle, ctypes.byref(n), ctypes.byref(b),
                                                          ctypes.byref(t)))
        return int(n.value), int(b.value), float(t.value)

    # ---- apply ---------------------------------------------------------------
    def _apply(self, out, state, transpose: bool, accumulate: bool):
        lib = _lib.lib()
        if _is_torch(out) or _is_torch(state):
            if not (_is_torch(out) and _is_torch(state)) or not out.is_cuda or not state.is_cuda:
                raise ArgumentError("device vectors must both be GPU tensors")
            prec = _torch_prec(out)
            if prec is None or state.dtype != out.dtype:
                raise ArgumentError("device vectors must share one of float64/complex128/float32/complex64")
            if not (out.is_contiguous() and state.is_contiguous()):
                raise ArgumentError("device vectors must be contiguous")
            _device_call(self.ctx, lib.edhip_apply, self._handle, prec, int(transpose), int(accumulate), state.numel(),
                         ctypes.c_void_p(state.data_ptr()), out.numel(), ctypes.c_void_p(out.data_ptr()))
            return out
        out_arr = out
        if not isinstance(out_arr, np.ndarray):
            raise ArgumentError("out must be a numpy array or a GPU tensor")
        prec = _lib.DTYPE_PREC.get(out_arr.dtype)
        if prec is None:
            raise ArgumentError(f"unsupported vector type {out_arr.dtype}")
        st = np.asarray(state)
        if np.iscomplexobj(st) and not np.iscomplexobj(out_arr):
            raise ArgumentError("InexactError: complex state into a real output vector")
        x = np.ascontiguousarray(st, dtype=out_arr.dtype)
        y = out_arr if out_arr.flags.c_contiguous else np.ascontiguousarray(out_arr)
        self.ctx.check(
            lib.edhip_apply_host(self._handle, prec, int(transpose), int(accumulate), x.size,
                                 x.ctypes.data_as(ctypes.c_void_p), y.size, y.ctypes.data_as(ctypes.c_void_p))
        )
        if y is not out_arr:
            out_arr[...] = y
        return out_arr

    def mul_rows(self, out, state, row_begin: int, row_end: int, transpose: bool = False):
        """One share of a row-sharded product (no counterpart in the
        reference): ``out[i] = (opr * state)[row_begin + i]`` for the rows
        ``[row_begin, row_end)``, with the whole ``state`` given (the input a
        GPU holds after an all-gather); ``transpose`` takes ``state * opr``.
        Vectors are GPU tensors of one dtype, or numpy arrays (copied through
        the device).  Every row equals the whole product's bitwise."""
        import torch

        lib = _lib.lib()
        host = not (_is_torch(out) and _is_torch(state))
        if host:
            if _is_torch(out) or _is_torch(state):
                raise ArgumentError("vectors must both be GPU tensors or both numpy arrays")
            if not isinstance(out, np.ndarray) or _lib.DTYPE_PREC.get(out.dtype) is None:
                raise ArgumentError("out must be a numpy array of float64/complex128/float32/complex64")
            st = np.asarray(state)
            if np.iscomplexobj(st) and not np.iscomplexobj(out):
                raise ArgumentError("InexactError: complex state into a real output vector")
            xs = torch.from_numpy(np.ascontiguousarray(st, dtype=out.dtype)).cuda()
            ys = torch.empty(out.size, dtype=xs.dtype, device=xs.device)
        else:
            if not (out.is_cuda and state.is_cuda) or state.dtype != out.dtype or _torch_prec(out) is None:
                raise ArgumentError("device vectors must be GPU tensors of one of float64/complex128/float32/complex64")
            if not (out.is_contiguous() and state.is_contiguous()):
                raise ArgumentError("device vectors must be contiguous")
            xs, ys = state, out
        _device_call(self.ctx, lib.edhip_apply_rows, self._handle, _torch_prec(ys), int(transpose), int(row_begin),
                     int(row_end), xs.numel(), ctypes.c_void_p(xs.data_ptr()), ys.numel(),
                     ctypes.c_void_p(ys.data_ptr()))
        if host:
            out[...] = ys.cpu().numpy()
        return out

    def __matmul__(self, state):
        """opr * state (operator_representation.jl:122-129): a new vector of the promoted type."""
        st = np.asarray(state)
        dt = np.result_type(st.dtype, self.dtype)
        if dt not in _lib.DTYPE_PREC:
            dt = np.dtype(np.complex128) if np.iscomplexobj(np.zeros(0, dt)) else np.dtype(np.float64)
        out = np.zeros(self.dimension, dtype=dt)
        return self._apply(out, st, False, True)

    __mul__ = __matmul__

    def __rmatmul__(self, state):
        """state * opr (operator_representation.jl:132-139)."""
        st = np.asarray(state)
        dt = np.result_type(st.dtype, self.dtype)
        if dt not in _lib.DTYPE_PREC:
            dt = np.dtype(np.complex128) if np.iscomplexobj(np.zeros(0, dt)) else np.dtype(np.float64)
        out = np.zeros(self.dimension, dtype=dt)
        return self._apply(out, st, True, True)

    __rmul__ = __rmatmul__

    # ---- iterators and elements -------------------------------------------------
    def _check_index(self, i):
        d = self.dimension
        if not (0 <= int(i) < d):
            raise BoundsError(f"index {i} out of bounds for dimension {d}")

    def _iterate(self, rows, transpose):
        rows = np.ascontiguousarray(np.atleast_1d(rows), dtype=np.int64)
        for r in rows:
            self._check_index(r)
        T = max(self._n_terms, 1)
        idx = np.empty(rows.size * T, dtype=np.int64)
        re = np.empty(rows.size * T)
        im = np.empty(rows.size * T)
        self.ctx.check(
            _lib.lib().edhip_iterate(self._handle, int(transpose), rows.size, ptr(rows, ctypes.c_int64),
                                     ptr(idx, ctypes.c_int64), ptr(re, ctypes.c_double), ptr(im, ctypes.c_double))
        )
        return idx.reshape(rows.size, T), (re + 1j * im).reshape(rows.size, T)

    def _pairs(self, irow, transpose):
        idx, val = self._iterate([irow], transpose)
        m = idx[0] != -2
        cast = complex if self.is_complex else (lambda z: z.real)
        return [(int(i), cast(v)) for i, v in zip(idx[0][m], val[0][m])]

    def get_row_iterator(self, irow):
        """Pairs (icol or -1, amplitude), duplicates kept (operator_representation.jl:58-79)."""
        return self._pairs(irow, False)

    def get_column_iterator(self, icol):
        return self._pairs(icol, True)

    def sparse(self, tol: float = RTOL_DEFAULT):
        """sparse(opr; tol): scipy CSC with the reference's merge/chop/sort."""
        import scipy.sparse as sp

        key = float(tol)
        if key not in self._sparse_cache:
            nnz = ctypes.c_int64()
            self.ctx.check(_lib.lib().edhip_sparse_build(self._handle, key, ctypes.byref(nnz)))
            n = self.dimension
            k = int(nnz.value)
            colptr = np.empty(n + 1, np.int64)
            rowval = np.empty(max(k, 1), np.int64)
            re = np.empty(max(k, 1))
            im = np.empty(max(k, 1))
            self.ctx.check(
                _lib.lib().edhip_sparse_export(self._handle, ptr(colptr, ctypes.c_int64), ptr(rowval, ctypes.c_int64),
                                               ptr(re, ctypes.c_double), ptr(im, ctypes.c_double))
            )
            data = (re[:k] + 1j * im[:k]) if self.is_complex else re[:k].copy()
            self._sparse_cache[key] = sp.csc_matrix((data, rowval[:k].copy(), colptr), shape=(n, n))
        return self._sparse_cache[key]

    sparse_serial = sparse
    sparse_parallel = sparse

    def Matrix(self):
        """Matrix(opr) (abstract_operator_representation.jl:121-132): the dense
        sum of the column iterators, no chop, built on the device."""
        n = self.dimension
        cx = self.is_complex
        if n == 0:
            return np.zeros((0, 0), np.complex128 if cx else np.float64)
        flat = np.empty(n * n * (2 if cx else 1))
        L = _lib.lib()
        buf = ctypes.c_void_p()
        self.ctx.check(L.edhip_malloc(self.ctx.handle, flat.nbytes, ctypes.byref(buf)))
        try:
            self.ctx.check(L.edhip_matrix(self._handle, buf))
            self.ctx.check(L.edhip_memcpy(self.ctx.handle, flat.ctypes.data_as(ctypes.c_void_p), buf, flat.nbytes, 2))
        finally:
            self.ctx.check(L.edhip_free(self.ctx.handle, buf))
        m = flat.view(np.complex128) if cx else flat
        return m.reshape(n, n).T.copy()  # column-major on the device

    def _vectors(self, which, transpose, tol):
        """Merged, chopped, ascending (index, value) arrays of rows or columns."""
        which = np.ascontiguousarray(np.atleast_1d(which), dtype=np.int64)
        for i in which:
            self._check_index(i)
        slot = self._n_terms + 1
        cnt = np.empty(which.size, np.int64)
        idx = np.empty(which.size * slot, np.int64)
        re = np.empty(which.size * slot)
        im = np.empty(which.size * slot)
        self.ctx.check(_lib.lib().edhip_get_vectors(
            self._handle, int(transpose), which.size, ptr(which, ctypes.c_int64), float(tol), ptr(cnt, ctypes.c_int64),
            ptr(idx, ctypes.c_int64), ptr(re, ctypes.c_double), ptr(im, ctypes.c_double)))
        out = []
        for k in range(which.size):
            s = slice(k * slot, k * slot + int(cnt[k]))
            v = re[s] + 1j * im[s] if self.is_complex else re[s].copy()
            out.append((idx[s].copy(), v))
        return out

    def get_row(self, irow):
        """get_row(opr, irow) (abstract_operator_representation.jl:217-229): the
        merged row iterator with choptol!(rtoldefault); a 1 x dim sparse row."""
        import scipy.sparse as sp

        (i, v), = self._vectors([irow], False, RTOL_DEFAULT)
        n = self.dimension
        return sp.csr_matrix((v, i, np.array([0, i.size])), shape=(1, n))

    def get_column(self, icol):
        """get_column(opr, icol) (:232-243): the merged column iterator with
        choptol!(rtoldefault); a dim x 1 sparse column."""
        import scipy.sparse as sp

        (i, v), = self._vectors([icol], True, RTOL_DEFAULT)
        n = self.dimension
        return sp.csc_matrix((v, i, np.array([0, i.size])), shape=(n, 1))

    def get_element(self, irow, icol):
        """get_element(opr, irow, icol) (operator_representation.jl:109-119,
        reduced_operator_representation.jl:120-138): the sum of column icol's
        iterator entries at row irow."""
        self._check_index(irow)
        (i, v), = self._vectors([icol], True, -1.0)
        hit = np.nonzero(i == int(irow))[0]
        zero = 0j if self.is_complex else 0.0
        return v[hit[0]] if hit.size else zero

    def __getitem__(self, key):
        if isinstance(key, tuple) and len(key) == 2:
            i, j = key
            if isinstance(i, slice) and isinstance(j, slice):
                return self.sparse()
            if isinstance(i, slice):
                return self.get_column(j)
            if isinstance(j, slice):
                return self.get_row(i)
            return self.get_element(i, j)
        raise TypeError("index with [i, j]")

    # ---- operator algebra on representations ---------------------------------------
    def _combine(self, other, f):
        if not isinstance(other, AbstractOperatorRepresentation) or other.space is not self.space:
            raise ArgumentError("representations must share the same space")
        return represent(self.space, f(self.operator, other.operator))

    def __add__(self, other):
        return self._combine(other, lambda a, b: a + b)

    def __sub__(self, other):
        return self._combine(other, lambda a, b: a - b)

    def __neg__(self):
        return represent(self.space, -1 * self.operator)

    def __eq__(self, other):
        return (
            isinstance(other, AbstractOperatorRepresentation)
            and other.space is self.space
            and self.operator == other.operator
        )

    __hash__ = object.__hash__


class OperatorRepresentation(AbstractOperatorRepresentation):
    """operator_representation.jl:13-26."""

    @property
    def hilbert_space_representation(self):
        return self.space

    def __repr__(self):
        return f"OperatorRepresentation(dim={self.dimension}, terms={self._n_terms})"


class ReducedOperatorRepresentation(AbstractOperatorRepresentation):
    """reduced_operator_representation.jl:15-30; its scalar type is complex."""

    @property
    def reduced_hilbert_space_representation(self):
        return self.space

    def __repr__(self):
        return f"ReducedOperatorRepresentation(dim={self.dimension}, terms={self._n_terms})"


# ---------------------------------------------------------------------------
# Julia-named free functions
# ---------------------------------------------------------------------------


def _split(a, b):
    if isinstance(a, AbstractOperatorRepresentation):
        return a, b, False
    if isinstance(b, AbstractOperatorRepresentation):
        return b, a, True
    raise ArgumentError("apply_ needs an operator representation")


def apply_(out, a, b):
    """apply!(out, opr, state) (out += opr*state) or apply!(out, state, opr)
    (out += state*opr): abstract_operator_representation.jl:260-285."""
    opr, state, transpose = _split(a, b)
    return opr._apply(out, state, transpose, True)


apply_serial_ = apply_
apply_parallel_ = apply_


def mul_(out, a, b):
    """mul!(out, opr, state): zero-fill then apply! (abstract_operator_representation.jl:110-118)."""
    opr, state, transpose = _split(a, b)
    return opr._apply(out, state, transpose, False)


def sparse(opr, tol: float = RTOL_DEFAULT):
    return opr.sparse(tol)


def Matrix(opr):
    return opr.Matrix()


def get_row_iterator(opr, irow):
    return opr.get_row_iterator(irow)


def get_column_iterator(opr, icol):
    return opr.get_column_iterator(icol)


def get_element(opr, irow, icol):
    return opr.get_element(irow, icol)


def get_row(opr, irow):
    return opr.get_row(irow)


def get_column(opr, icol):
    return opr.get_column(icol)


def dimension(x):
    return x.dimension
