"""ctypes binding of libedhip.so (the C ABI declared in include/edhip.h).

The shared library is built in-tree next to this file.  Nothing here falls
back to another implementation: if the library or a GPU is missing, calls
raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libedhip.so")

EDHIP_OK = 0
EDHIP_ERR_ARGUMENT = 1
EDHIP_ERR_DIMENSION = 2
EDHIP_ERR_MEMORY = 3
EDHIP_ERR_DEVICE = 4
EDHIP_ERR_KEY = 5
EDHIP_ERR_BOUNDS = 6
EDHIP_ERR_NOCONV = 7

F64, C128, F32, C64 = 0, 1, 2, 3
LANCZOS_FP64, LANCZOS_MIXED = 0, 1

DTYPE_PREC = {
    np.dtype(np.float64): F64,
    np.dtype(np.complex128): C128,
    np.dtype(np.float32): F32,
    np.dtype(np.complex64): C64,
}


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


class BoundsError(IndexError):
    """Julia's BoundsError."""


class NotConvergedError(RuntimeError):
    """An iterative solver reached its iteration limit."""


class DeviceError(RuntimeError):
    """A HIP runtime failure (including: no GPU present)."""


_EXC = {
    EDHIP_ERR_ARGUMENT: ArgumentError,
    EDHIP_ERR_DIMENSION: DimensionMismatch,
    EDHIP_ERR_MEMORY: MemoryError,
    EDHIP_ERR_DEVICE: DeviceError,
    EDHIP_ERR_KEY: KeyError,
    EDHIP_ERR_BOUNDS: BoundsError,
    EDHIP_ERR_NOCONV: NotConvergedError,
}


class LanczosOpts(ctypes.Structure):
    _fields_ = [
        ("nev", ctypes.c_int),
        ("max_basis", ctypes.c_int),
        ("max_restarts", ctypes.c_int),
        ("mode", ctypes.c_int),
        ("want_vectors", ctypes.c_int),
        ("refine_basis", ctypes.c_int),
        ("max_refine", ctypes.c_int),
        ("check_every", ctypes.c_int),
        ("criterion", ctypes.c_int),
        ("reorth", ctypes.c_int),
        ("seed", ctypes.c_uint64),
        ("tol", ctypes.c_double),
        ("tol_fp32", ctypes.c_double),
        ("assemble", ctypes.c_int),
    ]


class LanczosStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("restarts", ctypes.c_int),
        ("matvecs", ctypes.c_int),
        ("matvecs_fp64", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("second_passes", ctypes.c_int),
        ("residual", ctypes.c_double * 16),
        ("unrefined", ctypes.c_double * 16),
        ("t_total", ctypes.c_double),
        ("t_matvec", ctypes.c_double),
        ("t_refine", ctypes.c_double),
        ("t_alloc", ctypes.c_double),
        ("t_ortho", ctypes.c_double),
        ("ortho_bytes", ctypes.c_double),
        ("full_reorth_steps", ctypes.c_int),
        ("assembled", ctypes.c_int),
    ]


class MinresStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("residual", ctypes.c_double),
                ("assembled", ctypes.c_int)]


# name -> (restype, argtypes); every symbol declared in include/edhip.h
_vp = ctypes.c_void_p
_i32, _i64, _u64, _dbl = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
_P = ctypes.POINTER
SIGNATURES = {
    "edhip_create": (_i32, [_i32, _P(_vp)]),
    "edhip_destroy": (None, [_vp]),
    "edhip_last_error": (ctypes.c_char_p, [_vp]),
    "edhip_set_stream": (_i32, [_vp, _vp]),
    "edhip_get_stream": (_i32, [_vp, _P(_vp)]),
    "edhip_debug_fail_allocs": (_i32, [_i32, _i32]),
    "edhip_debug_asm_shape": (_i32, [_i32, _i32]),
    "edhip_debug_block_order": (_i32, [_i32]),
    "edhip_synchronize": (_i32, [_vp]),
    "edhip_malloc": (_i32, [_vp, _u64, _P(_vp)]),
    "edhip_free": (_i32, [_vp, _vp]),
    "edhip_release_workspace": (_i32, [_vp]),
    "edhip_memcpy": (_i32, [_vp, _vp, _vp, _u64, _i32]),
    "edhip_space_sector": (_i32, [_vp, _i32, _i32, _P(ctypes.c_int32), _P(_vp)]),
    "edhip_space_from_list": (_i32, [_vp, _i32, _i64, _P(_u64), _P(_vp)]),
    "edhip_space_sector_qn": (_i32, [_vp, _i32, _P(_i32), _P(_i32), _i32, _P(_i32), _i32, _P(_i32), _P(_vp)]),
    "edhip_space_reduce": (_i32, [_vp, _vp, _i32, _P(ctypes.c_int32), _P(_dbl), _P(_dbl), _dbl, _P(_vp)]),
    "edhip_space_reduce_ex": (_i32, [_vp, _vp, _i32, _P(ctypes.c_int32), _P(ctypes.c_uint64), _P(_dbl), _P(_dbl), _dbl,
                                     _P(_vp)]),
    "edhip_space_destroy": (None, [_vp]),
    "edhip_space_dim": (_i32, [_vp, _P(_i64)]),
    "edhip_space_copy_basis": (_i32, [_vp, _P(_u64)]),
    "edhip_space_device_basis": (_i32, [_vp, _P(_vp)]),
    "edhip_space_copy_orbit": (_i32, [_vp, _P(ctypes.c_uint16)]),
    "edhip_space_lookup": (_i32, [_vp, _i64, _P(_u64), _P(_i64)]),
    "edhip_space_export_mapping": (_i32, [_vp, _P(_i64), _P(_dbl), _P(_dbl)]),
    "edhip_space_vector_reduce": (_i32, [_vp, _i32, _i64, _vp, _i64, _vp]),
    "edhip_space_vector_unreduce": (_i32, [_vp, _i32, _i64, _vp, _i64, _vp]),
    "edhip_space_info": (_i32, [_vp, _P(_i32), _P(_i32), _P(_i32)]),
    "edhip_op_create": (_i32, [_vp, _vp, _i64, _P(_u64), _P(_u64), _P(_u64), _P(_dbl), _P(_dbl), _P(_vp)]),
    "edhip_op_destroy": (None, [_vp]),
    "edhip_op_is_complex": (_i32, [_vp, _P(_i32)]),
    "edhip_apply": (_i32, [_vp, _i32, _i32, _i32, _i64, _vp, _i64, _vp]),
    "edhip_apply_host": (_i32, [_vp, _i32, _i32, _i32, _i64, _vp, _i64, _vp]),
    "edhip_op_count_images": (_i32, [_vp, _P(_i64)]),
    "edhip_op_set_strategy": (_i32, [_vp, _i32, _i32]),
    "edhip_op_last_strategy": (_i32, [_vp, _P(_i32), _P(_i32)]),
    "edhip_apply_rows": (_i32, [_vp, _i32, _i32, _i64, _i64, _i64, _vp, _i64, _vp]),
    "edhip_op_assembled_info": (_i32, [_vp, _P(_u64), _P(_u64), _P(_dbl)]),
    "edhip_iterate": (_i32, [_vp, _i32, _i64, _P(_i64), _P(_i64), _P(_dbl), _P(_dbl)]),
    "edhip_sparse_build": (_i32, [_vp, _dbl, _P(_i64)]),
    "edhip_sparse_export": (_i32, [_vp, _P(_i64), _P(_i64), _P(_dbl), _P(_dbl)]),
    "edhip_get_vectors": (_i32, [_vp, _i32, _i64, _P(_i64), _dbl, _P(_i64), _P(_i64), _P(_dbl), _P(_dbl)]),
    "edhip_matrix": (_i32, [_vp, _vp]),
    "edhip_lanczos_default_opts": (None, [_P(LanczosOpts)]),
    "edhip_lanczos": (_i32, [_vp, _P(LanczosOpts), _P(_dbl), _vp, _P(LanczosStats)]),
    "edhip_term_expect": (_i32, [_vp, _i32, _vp, _vp, _P(_dbl)]),
    "edhip_term_expect_noalloc": (_i32, [_vp, _i32, _vp, _vp, _P(_dbl)]),
    "edhip_minres": (_i32, [_vp, _i32, _dbl, _i32, _vp, _vp, _vp, _dbl, _i32, _P(MinresStats)]),
    "edhip_vec_dot": (_i32, [_vp, _i32, _i64, _vp, _vp, _P(_dbl)]),
    "edhip_vec_gaussian": (_i32, [_vp, _i32, _i64, _u64, _vp]),
    "edhip_vec_convert": (_i32, [_vp, _i32, _i32, _i64, _vp, _vp]),
    "edhip_event_record": (_i32, [_vp, _i32]),
    "edhip_event_elapsed": (_i32, [_vp, _i32, _i32, _P(_dbl)]),
    "edhip_set_kernel_timing": (_i32, [_vp, _i32]),
    "edhip_op_kernel_time": (_i32, [_vp, _i32, _P(_dbl), _P(_i64)]),
}

_lib = None


def lib():
    """Load libedhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        # One HIP runtime per process: when PyTorch is present, load its
        # libamdhip64 first so that libedhip binds to the same runtime (both
        # carry the soname libamdhip64.so.7) and device pointers are shared.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class Context:
    """One edhip context (device + stream)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().edhip_create(int(device), ctypes.byref(h))
        if rc != EDHIP_OK:
            raise _EXC.get(rc, DeviceError)(f"edhip_create(device={device}) failed with status {rc}")
        self.handle = h
        self.device = int(device)

    def check(self, rc):
        if rc != EDHIP_OK:
            msg = lib().edhip_last_error(self.handle)
            msg = msg.decode() if msg else ""
            raise _EXC.get(rc, DeviceError)(msg)

    def synchronize(self):
        self.check(lib().edhip_synchronize(self.handle))

    def set_stream(self, stream_ptr):
        self.check(lib().edhip_set_stream(self.handle, ctypes.c_void_p(stream_ptr)))
        self._stream = None

    def stream_ptr(self) -> int:
        s = ctypes.c_void_p()
        self.check(lib().edhip_get_stream(self.handle, ctypes.byref(s)))
        return s.value or 0

    def order_after(self, torch_stream):
        """Run the context's next calls after the work queued so far on
        ``torch_stream``, with no host synchronisation: a non-default torch
        stream becomes the context's stream (one queue); for the default
        (null) stream, which the library cannot adopt, the context's own
        stream waits on an event of it.  Returns the torch stream the calls
        run on."""
        import torch

        ptr = torch_stream.cuda_stream
        if ptr:
            if getattr(self, "_stream", None) != ptr:
                self.set_stream(ptr)
                self._stream = ptr
            return torch_stream
        if getattr(self, "_stream", None) is not None:
            self.set_stream(0)
        own = torch.cuda.ExternalStream(self.stream_ptr(), device=torch_stream.device)
        own.wait_stream(torch_stream)
        return own

    def __del__(self):
        try:
            if getattr(self, "handle", None) and _lib is not None:
                _lib.edhip_destroy(self.handle)
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("EDHIP_DEVICE", "0")))
    return _default_ctx
