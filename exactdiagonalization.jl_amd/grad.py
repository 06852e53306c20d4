"""Differentiation of ground-state quantities with respect to couplings.

No counterpart exists in the reference; this is the backward pass of the
ground-state energy and of a correlator, expressed on top of the reference's
operator/representation API:

* H(theta) = sum_k c_k(theta) O_k for fixed operators O_k (e.g. per bond
  K_b = SxSx + SySy and Z_b = SzSz with c = (J_b, J_b Delta_b));
* Hellmann-Feynman: dE0/dtheta = sum_k dc_k/dtheta <psi|O_k|psi>, all <O_k>
  from one fused device pass over the basis (edhip_term_expect);
* correlator <O>: d<O>/dtheta = -2 Re sum_k dc_k/dtheta <lambda|O_k|psi> with
  (H - E0) lambda = P_perp (O - <O>) psi solved by MINRES on the device.

``GroundStateEnergy`` / ``GroundStateCorrelator`` are torch.autograd.Function
objects (the Python counterpart of a ChainRulesCore.rrule); their backward
recomputes psi by a second Lanczos pass instead of keeping it.
"""
from __future__ import annotations

from typing import Callable, Sequence

import numpy as np

from . import solve
from ._lib import ArgumentError
from .operators import NullOperator, PureOperator, SumOperator, simplify, terms_arrays
from .representation import (
    DirectProductOperation,
    GlobalBitFlip,
    ReducedHilbertSpaceRepresentation,
    SitePermutation,
    _perm_and_flip,
    represent,
)


def _key(t: PureOperator):
    return (t.bitmask, t.bitrow, t.bitcol)


def _term_map(op, perm_bits=None, xor=0):
    """{(mask, row, col): amplitude} of a simplified operator, after the bit
    permutation perm_bits (bit i -> bit perm_bits[i]) and the global flip
    (xor != 0: every bit of the mask inverted in row and column, the action
    of GlobalBitFlip, bitflipsymmetry.jl:23-35, on a PureOperator)."""

    def pb(w):
        if perm_bits is None:
            return w
        out, i = 0, 0
        while w:
            if w & 1:
                out |= 1 << int(perm_bits[i])
            w >>= 1
            i += 1
        return out

    out = {}
    for t in ([] if isinstance(op, NullOperator) else op.terms):
        m, r, c = t.bitmask, t.bitrow, t.bitcol
        if xor:
            r, c = m & ~r, m & ~c
        out[(pb(m), pb(r), pb(c))] = complex(t.amplitude)
    return out


def _check_invariant(space, operators, tol=1e-12):
    """On a symmetry-reduced space the representation applies an operator to
    the representative of each row only (reduced_operator_representation.jl:
    57-85), which is the operator's action on the symmetric states only when
    it commutes with every group element; a coupling operator that does not
    would give silently wrong <psi|O_k|psi>.  Raises ArgumentError."""
    if not isinstance(space, ReducedHilbertSpaceRepresentation):
        return
    hs = space.hilbert_space
    elems = []
    for g, _ in space.symops_and_amplitudes:
        if not isinstance(g, (SitePermutation, GlobalBitFlip, DirectProductOperation)):
            raise ArgumentError(f"unsupported symmetry operation {g!r}")
        elems.append(_perm_and_flip(hs, g))
    for k, o in enumerate(operators):
        base = _term_map(o)
        for perm, xor in elems:
            img = _term_map(o, perm, xor)
            if img.keys() != base.keys() or any(abs(img[q] - base[q]) > tol for q in base):
                raise ArgumentError(
                    f"operator {k} is not invariant under the symmetry group of the reduced space: its "
                    "expectation in the reduced basis is not that of the operator")


class LinearCouplings:
    """H(c) = sum_k c_k O_k over a fixed list of operators on one space."""

    def __init__(self, space, operators: Sequence):
        self.space = space
        self.operators = [simplify(o) for o in operators]
        _check_invariant(space, self.operators)
        keys = {}
        for o in self.operators:
            for t in ([] if isinstance(o, NullOperator) else o.terms):
                keys.setdefault(_key(t), len(keys))
        self.keys = list(keys.keys())
        A = np.zeros((len(self.operators), len(self.keys)), dtype=np.complex128)
        for k, o in enumerate(self.operators):
            for t in ([] if isinstance(o, NullOperator) else o.terms):
                A[k, keys[_key(t)]] += complex(t.amplitude)
        self.A = A
        # structure operator: every distinct term with unit amplitude
        self.structure = represent(space, SumOperator([PureOperator(m, r, c, 1.0) for (m, r, c) in self.keys]))

    def hamiltonian(self, c) -> object:
        c = np.asarray(c, dtype=np.float64)
        op = NullOperator()
        terms = []
        for ck, o in zip(c, self.operators):
            if isinstance(o, NullOperator):
                continue
            terms.extend(PureOperator(t.bitmask, t.bitrow, t.bitcol, ck * t.amplitude) for t in o.terms)
        if terms:
            op = simplify(SumOperator(terms))
        return represent(self.space, op)

    def expectations(self, phi, psi) -> np.ndarray:
        """<phi|O_k|psi> for every k from one fused pass."""
        g = solve.term_expectations(self.structure, phi, psi)
        return self.A @ g


def ground_state(model: LinearCouplings, c, *, tol=1e-11, seed=20241202, max_basis=160, return_vector=True):
    H = model.hamiltonian(c)
    r = solve.eigs(H, 1, tol=tol, seed=seed, max_basis=max_basis, return_vectors=return_vector)
    return H, r


def energy_and_grad(model: LinearCouplings, c, dc_dtheta, **kw):
    """E0 and dE0/dtheta (Hellmann-Feynman) for c = c(theta) with Jacobian
    dc_dtheta (n_c x n_theta)."""
    H, r = ground_state(model, c, **kw)
    psi = r.vectors[0]
    ex = model.expectations(psi, psi)
    grad = np.asarray(dc_dtheta).T @ ex.real
    return float(r.values[0]), grad


def correlator_and_grad(model: LinearCouplings, c, dc_dtheta, observable_op, *, minres_tol=1e-12, **kw):
    """<O> in the ground state and d<O>/dtheta via the MINRES adjoint."""
    import torch

    _check_invariant(model.space, [simplify(observable_op)])
    H, r = ground_state(model, c, **kw)
    E0 = float(r.values[0])
    psi = r.vectors[0].contiguous()
    O = represent(model.space, observable_op)
    Opsi = torch.zeros_like(psi)
    from .representation import apply_

    apply_(Opsi, O, psi)
    expO = torch.vdot(psi, Opsi).real.item()
    rhs = (Opsi - expO * psi).contiguous()
    lam, iters, res = solve.minres(H, rhs, shift=E0, deflate=psi, tol=minres_tol)
    ex = model.expectations(lam, psi)
    grad = -2.0 * (np.asarray(dc_dtheta).T @ ex.real)
    return expO, grad, {"minres_iterations": iters, "minres_residual": res, "E0": E0, "psi": psi, "lambda": lam}


def _torch_function():
    import torch

    class GroundStateEnergy(torch.autograd.Function):
        """E0(theta) with a Hellmann-Feynman backward.

        forward(theta, model, coupling_fn): coupling_fn(theta_np) -> (c, dc/dtheta).
        """

        @staticmethod
        def forward(ctx, theta, model, coupling_fn, opts=None):
            opts = opts or {}
            th = theta.detach().cpu().numpy()
            c, _ = coupling_fn(th)
            H = model.hamiltonian(c)
            r = solve.eigs(H, 1, return_vectors=False, **{k: v for k, v in opts.items() if k != "return_vectors"})
            ctx.model, ctx.coupling_fn, ctx.opts = model, coupling_fn, opts
            ctx.save_for_backward(theta)
            return theta.new_tensor(float(r.values[0]))

        @staticmethod
        def backward(ctx, grad_out):
            (theta,) = ctx.saved_tensors
            th = theta.detach().cpu().numpy()
            c, J = ctx.coupling_fn(th)
            # second Lanczos pass recomputes psi (same seed, same start vector)
            _, g = energy_and_grad(ctx.model, c, J, **ctx.opts)
            return grad_out * theta.new_tensor(g), None, None, None

    class GroundStateCorrelator(torch.autograd.Function):
        """<O>(theta) in the ground state with a MINRES-adjoint backward."""

        @staticmethod
        def forward(ctx, theta, model, coupling_fn, observable_op, opts=None):
            opts = opts or {}
            _check_invariant(model.space, [simplify(observable_op)])
            th = theta.detach().cpu().numpy()
            c, _ = coupling_fn(th)
            H, r = ground_state(model, c, **opts)
            psi = r.vectors[0]
            O = represent(model.space, observable_op)
            from .representation import apply_

            Opsi = torch.zeros_like(psi)
            apply_(Opsi, O, psi)
            val = torch.vdot(psi, Opsi).real.item()
            ctx.model, ctx.coupling_fn, ctx.obs, ctx.opts = model, coupling_fn, observable_op, opts
            ctx.save_for_backward(theta)
            return theta.new_tensor(val)

        @staticmethod
        def backward(ctx, grad_out):
            (theta,) = ctx.saved_tensors
            th = theta.detach().cpu().numpy()
            c, J = ctx.coupling_fn(th)
            _, g, _ = correlator_and_grad(ctx.model, c, J, ctx.obs, **ctx.opts)
            return grad_out * theta.new_tensor(g), None, None, None, None

    return GroundStateEnergy, GroundStateCorrelator


def xxz_model(hs, space, bonds):
    """Random-bond XXZ: H = sum_b J_b (SxSx + SySy + Delta_b SzSz) in S units.
    Returns (model, coupling_fn) with theta = (J_1..J_B, Delta_1..Delta_B)."""
    from .operators import pauli_matrix

    K, Z = [], []
    for (i, j) in bonds:
        sx = pauli_matrix(hs, i, "x") * pauli_matrix(hs, j, "x")
        sy = pauli_matrix(hs, i, "y") * pauli_matrix(hs, j, "y")
        sz = pauli_matrix(hs, i, "z") * pauli_matrix(hs, j, "z")
        K.append(simplify(0.25 * (sx + sy)))
        Z.append(simplify(0.25 * sz))
    model = LinearCouplings(space, K + Z)
    B = len(bonds)

    def coupling_fn(theta):
        J, D = theta[:B], theta[B:]
        c = np.concatenate([J, J * D])
        jac = np.zeros((2 * B, 2 * B))
        jac[np.arange(B), np.arange(B)] = 1.0
        jac[B + np.arange(B), np.arange(B)] = D
        jac[B + np.arange(B), B + np.arange(B)] = J
        return c, jac

    return model, coupling_fn
