"""edhip: MI355X-native Hamiltonian application on Sz-sector and
symmetry-reduced bases, behind the ExactDiagonalization.jl (v0.14.3) API.

Host-side Python mirror of the reference's Julia interface for this path;
every numerical step runs in the HIP kernels of ``libedhip.so``
(csrc/*.hip, C ABI in include/edhip.h).
"""
from ._lib import (
    ArgumentError,
    BoundsError,
    Context,
    DeviceError,
    DimensionMismatch,
    NotConvergedError,
    default_context,
)
from .hilbert import HilbertSpace, HilbertSpaceSector, Site, State, basespace
from .lattice import SquareSpaceIrrep, TranslationIrrep, chain_bonds, point_perm, square_bonds, translation_group, translation_perm
from .operators import (
    NullOperator,
    PureOperator,
    SumOperator,
    heisenberg_bond,
    pauli_matrix,
    pure_operator,
    simplify,
    spin_half_system,
    terms_arrays,
)
from .representation import (
    HilbertSpaceRepresentation,
    Matrix,
    OperatorRepresentation,
    ReducedHilbertSpaceRepresentation,
    ReducedOperatorRepresentation,
    SitePermutation,
    apply_,
    apply_parallel_,
    apply_serial_,
    bit_permutation,
    dimension,
    get_column,
    get_column_iterator,
    get_element,
    get_row,
    get_row_iterator,
    isinvariant,
    mul_,
    represent,
    represent_array,
    sparse,
    symmetry_apply,
    symmetry_reduce,
    symmetry_reduce_parallel,
    symmetry_reduce_serial,
    symmetry_unreduce,
    GlobalBitFlip,
    DirectProductOperation,
    make_product_irrep,
)
from .solve import EigsResult, eigs, minres, release_workspace, term_expectations

__all__ = [n for n in dir() if not n.startswith("_")]
