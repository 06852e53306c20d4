"""Eigensolver and linear solver on the device.

The reference has no eigensolver of its own: its callers run
``Arpack.eigs(opr; nev, which=:SR)`` through ``mul!``
(docs/src/examples/spinhalf.md:26).  ``eigs`` here is the thick-restart
Lanczos of libedhip (FP64, or FP32 Krylov storage with an FP64 refinement),
with the tridiagonal/arrowhead eigenproblem solved on the host.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError
from .representation import _device_call


@dataclass
class EigsResult:
    values: np.ndarray
    vectors: object  # torch tensor (nev, dim) on the GPU, or None
    iterations: int
    restarts: int
    matvecs: int
    matvecs_fp64: int
    converged: bool
    residuals: np.ndarray
    unrefined: np.ndarray
    t_total: float
    t_matvec: float
    t_refine: float
    t_alloc: float = 0.0
    second_passes: int = 0
    t_ortho: float = 0.0
    ortho_bytes: float = 0.0
    full_reorth_steps: int = 0
    assembled: bool = False


def eigs(opr, nev: int = 1, which: str = "SR", *, mode: str = "fp64", tol: float = 1e-10, max_basis: int = 120,
         max_restarts: int = 50, seed: int = 20241202, return_vectors: bool = False, refine_basis: int = 12,
         max_refine: int = 20, check_every: int = 5, criterion: str = "residual",
         reorth: str = "auto", tol_fp32: float = 0.0, assemble: str = "auto") -> EigsResult:
    """Lowest ``nev`` eigenvalues of a Hermitian operator representation.

    mode="fp64": FP64 (ComplexF64) vectors.  mode="mixed": FP32 (ComplexF32)
    Krylov storage and products, FP64 reductions, run to the eigenvalue
    criterion at max(tol, tol_fp32) (tol_fp32=0: max(tol, 1e-10)); then per
    eigenvalue an FP64 residual correction started from the FP32 Ritz vector
    (FP64 Rayleigh quotient and residual; more FP64 Lanczos steps only when
    the residual bound does not hold yet).  Start vector: SplitMix64(seed) +
    Box-Muller Gaussian.  ``t_total`` excludes the one-time allocation of the
    Krylov workspace, reported as ``t_alloc``.  For vectors over 2^22
    elements ``t_ortho`` is the device time of the orthogonalisation kernels
    against the basis and ``ortho_bytes`` their algorithmic bytes.

    criterion="residual": converged when ||H y - theta y|| <= tol*max(1,|theta|).
    criterion="eigenvalue": also when r^2/gap <= tol*max(1,|theta|) with theta
    stationary between checks (a bound on the eigenvalue error).

    reorth="adaptive": one Gram-Schmidt pass against the Krylov basis per step,
    a second when the first cancelled more than 99% of the vector's norm
    squared; reorth="twice": always two passes; reorth="partial": partial
    reorthogonalisation (Simon 1984) in the first cycle, the two latest
    vectors at every step and the whole basis only when the omega recurrence
    estimates an overlap above sqrt(eps) of the storage precision;
    reorth="auto": "partial" for vectors over 2^22 elements, else "adaptive".
    ``full_reorth_steps`` counts the steps orthogonalised against the whole
    basis.

    assemble="auto": with the operator's kernel strategy on auto, a reduced
    space over 2^22 states gets assembled rows for the time of the solve
    (every row's images found once, then streamed by each product) when they
    take at most 16 GiB and fit beside the workspace; ``assembled`` reports
    it.  assemble="off": never.
    """
    if which != "SR":
        raise ArgumentError("only which='SR' (smallest real) is supported")
    o = _lib.LanczosOpts()
    _lib.lib().edhip_lanczos_default_opts(ctypes.byref(o))
    o.nev = int(nev)
    o.max_basis = int(max_basis)
    o.max_restarts = int(max_restarts)
    o.mode = _lib.LANCZOS_MIXED if mode == "mixed" else _lib.LANCZOS_FP64
    o.want_vectors = 1 if return_vectors else 0
    o.refine_basis = int(refine_basis)
    o.max_refine = int(max_refine)
    o.check_every = int(check_every)
    o.criterion = 1 if criterion == "eigenvalue" else 0
    modes = {"auto": 0, "twice": 1, "partial": 2, "adaptive": 3}
    if reorth not in modes:
        raise ArgumentError("reorth must be 'auto', 'adaptive', 'partial' or 'twice'")
    o.reorth = modes[reorth]
    if assemble not in ("auto", "off"):
        raise ArgumentError("assemble must be 'auto' or 'off'")
    o.assemble = 0 if assemble == "auto" else 1
    o.seed = int(seed)
    o.tol = float(tol)
    o.tol_fp32 = float(tol_fp32)
    st = _lib.LanczosStats()
    vals = np.zeros(nev)
    vecs = None
    vptr = ctypes.c_void_p(0)
    if return_vectors:
        import torch

        dt = torch.complex128 if opr.is_complex else torch.float64
        vecs = torch.empty((nev, opr.dimension), dtype=dt, device=f"cuda:{opr.ctx.device}")
        vptr = ctypes.c_void_p(vecs.data_ptr())
        _device_call(opr.ctx, _lib.lib().edhip_lanczos, opr._handle, ctypes.byref(o), _lib.ptr(vals, ctypes.c_double), vptr,
                     ctypes.byref(st))
    else:
        opr.ctx.check(
            _lib.lib().edhip_lanczos(opr._handle, ctypes.byref(o), _lib.ptr(vals, ctypes.c_double), vptr, ctypes.byref(st))
        )
    return EigsResult(
        values=vals,
        vectors=vecs,
        iterations=st.iterations,
        restarts=st.restarts,
        matvecs=st.matvecs,
        matvecs_fp64=st.matvecs_fp64,
        converged=bool(st.converged),
        second_passes=st.second_passes,
        residuals=np.array(st.residual[:nev]),
        unrefined=np.array(st.unrefined[:nev]),
        t_total=st.t_total,
        t_matvec=st.t_matvec,
        t_refine=st.t_refine,
        t_alloc=st.t_alloc,
        t_ortho=st.t_ortho,
        ortho_bytes=st.ortho_bytes,
        full_reorth_steps=st.full_reorth_steps,
        assembled=bool(st.assembled),
    )


def release_workspace(ctx=None):
    """Free the device workspace that eigs and minres keep between calls."""
    ctx = ctx or _lib.default_context()
    ctx.check(_lib.lib().edhip_release_workspace(ctx.handle))


def minres(opr, b, *, shift: float = 0.0, deflate=None, tol: float = 1e-10, max_iter: int = 2000,
           stats: bool = False):
    """Solve P (H - shift) P x = P b with MINRES, P = 1 - sum d d^H over the
    rows of ``deflate`` (orthonormal).  b, deflate: GPU tensors (float64 or
    complex128).  Returns (x, iterations, residual), and whether the solve
    assembled the operator's rows when ``stats`` (it does, as eigs, on
    reduced spaces over 2^22 states with the kernel strategy on auto)."""
    import torch

    if not (torch.is_tensor(b) and b.is_cuda):
        raise ArgumentError("b must be a GPU tensor")
    prec = {torch.float64: _lib.F64, torch.complex128: _lib.C128}.get(b.dtype)
    if prec is None:
        raise ArgumentError("MINRES runs in float64 or complex128")
    x = torch.empty_like(b)
    nd = 0
    dptr = ctypes.c_void_p(0)
    if deflate is not None:
        d = deflate.reshape(-1, b.numel()).contiguous()
        if d.dtype != b.dtype:
            raise ArgumentError("deflation vectors must match b's dtype")
        nd = d.shape[0]
        dptr = ctypes.c_void_p(d.data_ptr())
    st = _lib.MinresStats()
    _device_call(opr.ctx, _lib.lib().edhip_minres, opr._handle, prec, float(shift), nd, dptr, ctypes.c_void_p(b.data_ptr()),
                 ctypes.c_void_p(x.data_ptr()), float(tol), int(max_iter), ctypes.byref(st))
    if stats:
        return x, st.iterations, st.residual, bool(st.assembled)
    return x, st.iterations, st.residual


def term_expectations(opr, phi, psi) -> np.ndarray:
    """Unit-amplitude bilinear forms <phi|T_t|psi> of every term t of the
    operator, in term order (one fused pass on the device)."""
    import torch

    prec = {torch.float64: _lib.F64, torch.complex128: _lib.C128, torch.float32: _lib.F32,
            torch.complex64: _lib.C64}.get(psi.dtype)
    if prec is None or phi.dtype != psi.dtype:
        raise ArgumentError("phi and psi must share a supported dtype")
    out = np.zeros(2 * max(opr._n_terms, 1))
    _device_call(opr.ctx, _lib.lib().edhip_term_expect, opr._handle, prec, ctypes.c_void_p(phi.data_ptr()),
                 ctypes.c_void_p(psi.data_ptr()), _lib.ptr(out, ctypes.c_double))
    return out[0::2][: opr._n_terms] + 1j * out[1::2][: opr._n_terms]
