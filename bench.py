"""Benchmark: symmetry-reduced Hamiltonian application on one MI355X.

Workload (default): J1-J2 Heisenberg model on the 6x6 periodic square
lattice (36 sites, J1 = 1, J2 = 0.5 on both diagonals, S.S units), Sz = 0,
reduced by the 36 translations at momentum (0, 0): 252,091,362
representatives.  One step = one matrix-free matvec y = H x over the whole
reduced space, vectors and basis resident in HBM.

Prints one JSON line.  ``value`` is the element throughput (Gelem/s) of the
FP32-storage matvec (the one the mixed-precision Lanczos runs); the FP64
matvec is reported beside it.  ``roofline`` prices the algorithmic bytes of
one matvec (basis word + own x + y per row, key + x per off-diagonal image)
over the matvec kernel's own launch time (HIP events around each launch, the
spread of x that precedes it excluded) against the 8 TB/s HBM peak.
``cpu_baseline`` times the oracle's restatement of the reference CPU path
(basis, symmetry_reduce, sparse_parallel on the host cores, ARPACK) on the
configurations that path can build; the 6x6 workload is beyond it.

``grad`` times the backward pass on the 4x6 random-bond XXZ model.

``time_to_E0`` times the FP32-mixed and the FP64 Lanczos to E0 on the
workload's operator.

``row_shares`` times one GPU's share of a product sharded by row ranges
over 8 GPUs (each eighth of the rows, the whole x resident).

``assembled_rows`` times the assembled-row strategy on the same operator:
the build of its entries, its FP32 and FP64 products and both Lanczos modes
on them.

Options: --steps K --warmup W --gpus 1 [--config 4x6|6x6] [--no-cpu]
[--no-grad] [--no-e0] [--no-assembled] [--no-shares] [--no-space-group].
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "6x6": dict(Lx=6, Ly=6, k=(0, 0), J2=0.5),
    "4x6": dict(Lx=4, Ly=6, k=(2, 0), J2=0.5),
}


def build_workload(ed, cfg):
    Lx, Ly = cfg["Lx"], cfg["Ly"]
    n = Lx * Ly
    hs, _ = ed.spin_half_system(n)
    j1, j2 = ed.square_bonds(Lx, Ly)
    terms = [ed.heisenberg_bond(hs, i, j, 1.0) for i, j in j1] + [ed.heisenberg_bond(hs, i, j, cfg["J2"]) for i, j in j2]
    H = ed.simplify(sum(terms, ed.NullOperator()))
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    t0 = time.perf_counter()
    rhsr = ed.symmetry_reduce(hsr, ed.TranslationIrrep((Lx, Ly), cfg["k"]))
    t_reduce = time.perf_counter() - t0
    opr = ed.represent(rhsr, H)
    return hs, H, hsr, rhsr, opr, t_reduce


def time_matvecs(ed, opr, prec, steps, warmup):
    import torch
    from edhip import _lib

    L = _lib.lib()
    ctx = opr.ctx
    D = opr.dimension
    dt = {_lib.F32: torch.float32, _lib.F64: torch.float64}[prec]
    g = torch.Generator(device="cuda").manual_seed(1234)
    x = torch.randn(D, dtype=dt, device="cuda", generator=g)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    xp, yp = ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr())
    for _ in range(warmup):
        ctx.check(L.edhip_apply(opr._handle, prec, 0, 0, D, xp, D, yp))
    ctx.synchronize()
    # timed region: HIP events on the library's stream, where the matvec runs
    ctx.check(L.edhip_event_record(ctx.handle, 0))
    for _ in range(steps):
        ctx.check(L.edhip_apply(opr._handle, prec, 0, 0, D, xp, D, yp))
    ctx.check(L.edhip_event_record(ctx.handle, 1))
    ms = ctypes.c_double()
    ctx.check(L.edhip_event_elapsed(ctx.handle, 0, 1, ctypes.byref(ms)))
    ctx.synchronize()
    checksum = float(y.double().sum().item())
    del x, y
    torch.cuda.empty_cache()
    return ms.value / steps, checksum


def kernel_ms(ed, opr, prec, steps):
    """Average duration of the matvec kernel launches alone (HIP events around
    each launch on the library's stream; the spread of x before it excluded)."""
    import torch
    from edhip import _lib

    L = _lib.lib()
    ctx = opr.ctx
    D = opr.dimension
    dt = {_lib.F32: torch.float32, _lib.F64: torch.float64}[prec]
    x = torch.randn(D, dtype=dt, device="cuda", generator=torch.Generator(device="cuda").manual_seed(99))
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    xp, yp = ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr())
    tot, n = ctypes.c_double(), ctypes.c_int64()
    ctx.check(L.edhip_op_kernel_time(opr._handle, 1, ctypes.byref(tot), ctypes.byref(n)))
    ctx.check(L.edhip_set_kernel_timing(ctx.handle, 1))
    try:
        for _ in range(steps):
            ctx.check(L.edhip_apply(opr._handle, prec, 0, 0, D, xp, D, yp))
        ctx.synchronize()
    finally:
        ctx.check(L.edhip_set_kernel_timing(ctx.handle, 0))
    ctx.check(L.edhip_op_kernel_time(opr._handle, 1, ctypes.byref(tot), ctypes.byref(n)))
    del x, y
    torch.cuda.empty_cache()
    return tot.value / max(n.value, 1)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ed):
    """The reference's CPU path, restated by the oracle, on the host cores:
    parent sector basis, symmetry_reduce (serial restatement of
    symmetry_reduce_generic.jl:22-106), sparse_parallel (OpenMP over column
    blocks, abstract_operator_representation.jl:172-204) and ARPACK
    (scipy eigsh, which='SA', the library Arpack.jl wraps) for the configs
    it can build: the 16-site chain, the 24-site chain at k = 3 (4 levels),
    the 4x6 J1-J2 lattice at (pi, 0), and the 4x6 random-bond XXZ sector.
    The 6x6 workload is out of its reach: its parent arrays hold 9.08e9
    words (>= 290 GB).  ``value``: elements applied per second along the
    4x6 J1-J2 path to E0, (nnz x ARPACK products) / (sparse build + eigsh)."""
    import oracle
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1))

    def perms_chars(Lx, Ly, k):
        irrep = ed.TranslationIrrep((Lx, Ly), k)
        hs, _ = ed.spin_half_system(Lx * Ly)
        ops = irrep.symops_and_amplitudes()
        return np.stack([ed.bit_permutation(hs, p) for p, _ in ops]), [a for _, a in ops]

    def run(name, n, H, group, nev):
        t0 = time.perf_counter()
        parent = oracle.sector_basis(n, 0)
        t_basis = time.perf_counter() - t0
        space, t_red = parent, 0.0
        if group is not None:
            t0 = time.perf_counter()
            space = oracle.symmetry_reduce(parent, n, *group)
            t_red = time.perf_counter() - t0
        o = oracle.Opr(space, ed.terms_arrays(H))
        t0 = time.perf_counter()
        colptr, rowval, nz = o.sparse_parallel(n_threads=threads)
        t_sparse = time.perf_counter() - t0
        A = sp.csc_matrix((nz, rowval, colptr), shape=(o.dim, o.dim))
        if np.all(A.data.imag == 0):
            A = A.real
        A = A.tocsr()
        count = [0]

        def mv(x):
            count[0] += 1
            return A @ x

        op = sla.LinearOperator(A.shape, matvec=mv, dtype=A.dtype)
        v0 = np.random.default_rng(20241202).standard_normal(o.dim).astype(A.dtype)
        t0 = time.perf_counter()
        w = np.sort(sla.eigsh(op, k=nev, which="SA", tol=1e-10, v0=v0, return_eigenvectors=False))
        t_eig = time.perf_counter() - t0
        return {"config": name, "dim": o.dim, "nnz": int(A.nnz), "basis_s": t_basis, "reduce_s": t_red,
                "sparse_s": t_sparse, "eigsh_s": t_eig, "arpack_products": count[0], "E": [float(v) for v in w]}

    def heis(hs, bonds, w, d=None):
        return ed.simplify(sum((ed.heisenberg_bond(hs, i, j, w[k], 1.0 if d is None else d[k])
                                for k, (i, j) in enumerate(bonds)), ed.NullOperator()))

    rows = []
    hs, _ = ed.spin_half_system(16)
    rows.append(run("chain16", 16, heis(hs, ed.chain_bonds(16), [1.0] * 16), None, 1))
    hs, _ = ed.spin_half_system(24)
    rows.append(run("chain24_k3", 24, heis(hs, ed.chain_bonds(24), [1.0] * 24), perms_chars(24, 1, (3, 0)), 4))
    j1, j2 = ed.square_bonds(4, 6)
    rows.append(run("j1j2_4x6_pi0", 24, heis(hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2)),
                    perms_chars(4, 6, (2, 0)), 1))
    rng = np.random.default_rng(2024)
    B = len(j1)
    J, Dl = rng.uniform(0.5, 1.5, B), rng.uniform(0.8, 1.2, B)
    rows.append(run("xxz_4x6_random", 24, heis(hs, j1, J, Dl), None, 1))
    c3 = rows[2]
    value = c3["nnz"] * c3["arpack_products"] / (c3["sparse_s"] + c3["eigsh_s"]) / 1e9
    total = sum(r["basis_s"] + r["reduce_s"] + r["sparse_s"] + r["eigsh_s"] for r in rows)
    return {
        "value": value,
        "unit": "Gelem/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "sample": (
            f"oracle restatement of the reference CPU path (basis, serial symmetry_reduce, OpenMP sparse_parallel "
            f"on {threads} threads, scipy ARPACK eigsh SA tol 1e-10) on the 16-site chain, the 24-site chain k=3 "
            f"(4 levels), 4x6 J1-J2 (pi,0) and 4x6 random XXZ: {total:.1f} s in all; value = 4x6 J1-J2 "
            f"nnz x ARPACK products / (sparse + eigsh seconds); the 6x6 k=(0,0) workload is N/A for this "
            f"algorithm (parent arrays over 9.08e9 words, >= 290 GB)"
        ),
        "configs": rows,
        "not_runnable": {"j1j2_6x6_k00": "parent-sector arrays of 9,075,135,300 words (>= 290 GB) exceed host memory"},
    }


def lanczos_times(ed, opr, tol_fp32=0.0, reorth="auto"):
    """Time to E0 at 1e-10 relative (eigenvalue criterion r^2/gap) for the
    FP32-storage mixed Lanczos (with the FP64 residual correction) and the
    FP64 Lanczos.  ``seconds`` runs from the start vector to E0; the one-time
    allocation of the Krylov workspace is excluded from both and reported as
    ``alloc_seconds`` (the first run allocates, the second reuses)."""
    out = {}
    for mode in ("fp64", "mixed"):
        r = ed.eigs(opr, 1, mode=mode, tol=1e-10, max_basis=100, max_restarts=20, criterion="eigenvalue",
                    tol_fp32=tol_fp32, reorth=reorth)
        out[mode] = {
            "E0": float(r.values[0]),
            "seconds": r.t_total,
            "alloc_seconds": r.t_alloc,
            "iterations": r.iterations,
            "matvecs": r.matvecs,
            "matvecs_fp64": r.matvecs_fp64,
            "matvec_seconds": r.t_matvec,
            "refine_seconds": r.t_refine,
            "other_seconds": r.t_total - r.t_matvec - r.t_refine,
            "reorth": reorth,
            "full_reorth_steps": r.full_reorth_steps,
            "reorth_second_passes": r.second_passes,
            # orthogonalisation kernels (dot products and updates against the
            # basis): device time from HIP events, algorithmic bytes over it
            "ortho_kernel_seconds": r.t_ortho,
            "ortho_GBps": r.ortho_bytes / r.t_ortho / 1e9 if r.t_ortho > 0 else None,
            "ortho_frac_of_8TBps": r.ortho_bytes / r.t_ortho / 8e12 if r.t_ortho > 0 else None,
            "residual": float(r.residuals[0]),
            "converged": r.converged,
        }
        if mode == "mixed":
            out[mode]["E0_fp32_phase"] = float(r.unrefined[0])
            out[mode]["tol_fp32_phase"] = max(1e-10, tol_fp32 if tol_fp32 > 0 else 1e-10)
    out["speedup_mixed_vs_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["rel_diff_E0"] = abs(out["mixed"]["E0"] - out["fp64"]["E0"]) / abs(out["fp64"]["E0"])
    return out


def assembled_times(ed, opr, elems, steps, warmup, e0=True):
    """The assembled-row strategy on the same workload: its one-time build
    (every row's images found once, kept as 4-byte entries), the FP32 and
    FP64 products that stream the entries, and the time to E0 of both
    Lanczos modes on them (basis of 80 vectors: the FP64 workspace and the
    entries then fit together; both modes converge before a restart).  The
    matrix-free product stays the headline; this is the alternative for
    repeated products when memory allows."""
    from edhip import _lib

    ed.release_workspace()  # the time-to-E0 workspace would leave no room for the entries
    opr.set_strategy("assembled")
    try:
        n_ent, n_bytes, t_build = opr.assembled_info()
        ms32, cs32 = time_matvecs(ed, opr, _lib.F32, steps, warmup)
        ms64, cs64 = time_matvecs(ed, opr, _lib.F64, steps, warmup)
        k32 = kernel_ms(ed, opr, _lib.F32, steps)
        k64 = kernel_ms(ed, opr, _lib.F64, steps)
        out = {
            "entries": n_ent,
            "device_bytes": n_bytes,
            "build_seconds": t_build,
            "f32": {"ms_per_step": ms32, "kernel_ms": k32, "Gelem_per_s": elems / (ms32 * 1e-3) / 1e9,
                    "checksum": cs32},
            "f64": {"ms_per_step": ms64, "kernel_ms": k64, "Gelem_per_s": elems / (ms64 * 1e-3) / 1e9,
                    "checksum": cs64},
            "kernels": "k_asm_scale (x / sqrt(N)) + k_asm_spmv (16 lanes per row, 8 entries per lane per batch)",
        }
        te = {}
        for mode in ("fp64", "mixed") if e0 else ():
            r = ed.eigs(opr, 1, mode=mode, tol=1e-10, max_basis=80, max_restarts=20, criterion="eigenvalue")
            te[mode] = {"E0": float(r.values[0]), "seconds": r.t_total, "alloc_seconds": r.t_alloc,
                        "matvecs": r.matvecs, "matvec_seconds": r.t_matvec, "converged": r.converged,
                        "strategy": opr.last_strategy()[0]}
        if e0:
            te["speedup_mixed_vs_fp64"] = te["fp64"]["seconds"] / te["mixed"]["seconds"]
            out["time_to_E0"] = te
    finally:
        opr.set_strategy()
        ed.release_workspace()
    return out


def share_times(ed, opr, n_shares=8, reps=3):
    """One GPU's share of a product sharded by contiguous row ranges over
    ``n_shares`` GPUs: the FP32 kernel time of each share's rows (the whole x
    resident, as after the all-gather), from the matrix-free kernel's own
    launch events.  The slowest share sets the sharded product's kernel time;
    the spread between shares is the load imbalance of equal row counts."""
    import torch
    from edhip import _lib

    L = _lib.lib()
    ctx = opr.ctx
    D = opr.dimension
    x = torch.randn(D, dtype=torch.float32, device="cuda", generator=torch.Generator(device="cuda").manual_seed(7))
    cuts = [D * k // n_shares for k in range(n_shares + 1)]
    ms = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        y = torch.empty(b - a, dtype=torch.float32, device="cuda")
        xp, yp = ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr())
        ctx.check(L.edhip_apply_rows(opr._handle, _lib.F32, 0, a, b, D, xp, b - a, yp))  # warm
        ctx.synchronize()
        tot, n = ctypes.c_double(), ctypes.c_int64()
        ctx.check(L.edhip_op_kernel_time(opr._handle, 1, ctypes.byref(tot), ctypes.byref(n)))
        ctx.check(L.edhip_set_kernel_timing(ctx.handle, 1))
        try:
            for _ in range(reps):
                ctx.check(L.edhip_apply_rows(opr._handle, _lib.F32, 0, a, b, D, xp, b - a, yp))
            ctx.synchronize()
        finally:
            ctx.check(L.edhip_set_kernel_timing(ctx.handle, 0))
        ctx.check(L.edhip_op_kernel_time(opr._handle, 1, ctypes.byref(tot), ctypes.byref(n)))
        ms.append(tot.value / max(n.value, 1))
        del y
    del x
    torch.cuda.empty_cache()
    return {
        "shares": n_shares,
        "rows_per_share": [b - a for a, b in zip(cuts[:-1], cuts[1:])],
        "f32_kernel_ms": ms,
        "max_ms": max(ms),
        "mean_ms": sum(ms) / len(ms),
        "sum_ms": sum(ms),
        "imbalance_max_over_mean": max(ms) / (sum(ms) / len(ms)),
        "note": "one GPU's share at its real shape (rows of D/8, full basis and x resident); "
                "a sharded product would add the all-gather of x, not measured (one GPU)",
    }


def small_config_times(ed):
    """The CPU baseline's configurations on the GPU, for a like-for-like view
    beside it: the 16-site chain, the 24-site chain at k = 3 (4 levels), the
    4x6 J1-J2 (pi,0) sector and the 4x6 random-bond XXZ sector (seed 2024).
    Per config: the device basis and reduction (wall seconds), one FP64
    product (HIP events, kernel launches only, averaged over 20), and the
    FP64 Lanczos to the lowest levels at tol 1e-10 (seconds from the start
    vector, the workspace excluded), with the levels."""
    import torch
    from edhip import _lib

    L = _lib.lib()

    def heis(hs, bonds, w, d=None):
        return ed.simplify(sum((ed.heisenberg_bond(hs, i, j, w[k], 1.0 if d is None else d[k])
                                for k, (i, j) in enumerate(bonds)), ed.NullOperator()))

    def run(name, n, H, irrep, nev):
        hs, _ = ed.spin_half_system(n)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        space = ed.represent(ed.HilbertSpaceSector(hs, 0))
        if irrep is not None:
            space = ed.symmetry_reduce(space, irrep)
        dim = space.dimension
        t_space = time.perf_counter() - t0
        opr = ed.represent(space, H)
        ctx = opr.ctx
        dt = torch.complex128 if opr.is_complex else torch.float64
        x = torch.randn(dim, dtype=torch.float64, device="cuda").to(dt)
        y = torch.empty_like(x)
        ed.mul_(y, opr, x)
        tot, nl = ctypes.c_double(), ctypes.c_int64()
        ctx.check(L.edhip_op_kernel_time(opr._handle, 1, ctypes.byref(tot), ctypes.byref(nl)))
        ctx.check(L.edhip_set_kernel_timing(ctx.handle, 1))
        try:
            for _ in range(20):
                ed.mul_(y, opr, x)
        finally:
            ctx.check(L.edhip_set_kernel_timing(ctx.handle, 0))
        ctx.check(L.edhip_op_kernel_time(opr._handle, 1, ctypes.byref(tot), ctypes.byref(nl)))
        r = ed.eigs(opr, nev, tol=1e-10, max_basis=min(100, dim), criterion="eigenvalue")
        return {"config": name, "dim": dim, "space_s": t_space, "matvec_kernel_ms": tot.value / max(nl.value, 1),
                "eigs_s": r.t_total, "matvecs": r.matvecs, "E": [float(v) for v in r.values]}

    rows = []
    hs, _ = ed.spin_half_system(16)
    rows.append(run("chain16", 16, heis(hs, ed.chain_bonds(16), [1.0] * 16), None, 1))
    hs, _ = ed.spin_half_system(24)
    rows.append(run("chain24_k3", 24, heis(hs, ed.chain_bonds(24), [1.0] * 24), ed.TranslationIrrep(24, 3), 4))
    j1, j2 = ed.square_bonds(4, 6)
    rows.append(run("j1j2_4x6_pi0", 24, heis(hs, j1 + j2, [1.0] * len(j1) + [0.5] * len(j2)),
                    ed.TranslationIrrep((4, 6), (2, 0)), 1))
    rng = np.random.default_rng(2024)
    B = len(j1)
    J, Dl = rng.uniform(0.5, 1.5, B), rng.uniform(0.8, 1.2, B)
    rows.append(run("xxz_4x6_random", 24, heis(hs, j1, J, Dl), None, 1))
    return rows


def grad_times(ed):
    """Backward pass on the 4x6 random-bond XXZ model (Sz = 0, 2,704,156
    states, 48 bonds, theta = (J_b, Delta_b), J_b ~ U(0.5,1.5), Delta_b ~
    U(0.8,1.2) from seed 2024): the ground state by FP64 Lanczos, the 96
    Hellmann-Feynman derivatives from one fused term pass, and
    d<Sz_0 Sz_12>/dtheta through the MINRES adjoint.  The fused pass is also
    expressed in matvec equivalents (its time over one FP64 matvec of H)."""
    import torch
    from edhip.grad import correlator_and_grad, ground_state, xxz_model

    Lx, Ly = 4, 6
    n = Lx * Ly
    hs, s = ed.spin_half_system(n)
    j1, _ = ed.square_bonds(Lx, Ly)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    model, fn = xxz_model(hs, hsr, j1)
    rng = np.random.default_rng(2024)
    B = len(j1)
    theta = np.concatenate([rng.uniform(0.5, 1.5, B), rng.uniform(0.8, 1.2, B)])
    c, J = fn(theta)
    t0 = time.perf_counter()
    H, r = ground_state(model, c, tol=1e-12)
    t_gs = time.perf_counter() - t0
    psi = r.vectors[0]
    model.expectations(psi, psi)  # warm-up (builds the structure operator's tables)
    reps = 20
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ex = model.expectations(psi, psi)
    t_bw = (time.perf_counter() - t0) / reps
    g = J.T @ ex.real
    # the fused term pass alone: its kernel and partial-sum reduction, HIP
    # events around each launch; algorithmic bytes = D * (word + phi + psi)
    # + images * (psi gather), the sector's index being rank arithmetic
    from edhip import _lib

    L = _lib.lib()
    st = model.structure
    ctx = st.ctx
    n_img = ctypes.c_int64()
    ctx.check(L.edhip_op_count_images(st._handle, ctypes.byref(n_img)))
    tot, nl = ctypes.c_double(), ctypes.c_int64()
    ctx.check(L.edhip_op_kernel_time(st._handle, 1, ctypes.byref(tot), ctypes.byref(nl)))
    ctx.check(L.edhip_set_kernel_timing(ctx.handle, 1))
    try:
        for _ in range(reps):
            model.expectations(psi, psi)
    finally:
        ctx.check(L.edhip_set_kernel_timing(ctx.handle, 0))
    ctx.check(L.edhip_op_kernel_time(st._handle, 1, ctypes.byref(tot), ctypes.byref(nl)))
    t_tk = tot.value / max(nl.value, 1) * 1e-3
    D = psi.numel()
    tk_bytes = D * (8 + 8 + 8) + n_img.value * 8
    y = torch.empty_like(psi)
    ed.mul_(y, H, psi)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ed.mul_(y, H, psi)
    torch.cuda.synchronize()  # device calls are stream-ordered: wait for the products
    t_mv = (time.perf_counter() - t0) / reps
    O = ed.simplify(0.25 * s(0, "z") * s(12, "z"))
    t0 = time.perf_counter()
    val, gc, info = correlator_and_grad(model, c, J, O, tol=1e-12, minres_tol=1e-10)
    t_corr = time.perf_counter() - t0
    return {
        "workload": "random-bond XXZ 4x6 torus, Sz=0, 2,704,156 states, 96 parameters (seed 2024)",
        "E0": float(r.values[0]),
        "ground_state_lanczos_s": t_gs,
        "lanczos_matvecs": int(r.matvecs),
        "energy_backward_ms": t_bw * 1e3,
        "fp64_matvec_ms": t_mv * 1e3,
        "energy_backward_matvec_equivalents": t_bw / t_mv,
        "term_pass_kernel_ms": t_tk * 1e3,
        "term_pass_images": int(n_img.value),
        "term_pass_alg_bytes": tk_bytes,
        "term_pass_GBps": tk_bytes / t_tk / 1e9,
        "term_pass_frac_of_8TBps": tk_bytes / t_tk / 8e12,
        "homogeneity_residual": float(abs(np.dot(theta[:B], g[:B]) - r.values[0])),
        "correlator": "Sz_0 Sz_12",
        "correlator_value": float(val),
        "correlator_grad_total_s": t_corr,
        "minres_iterations": int(info["minres_iterations"]),
        "minres_residual": float(info["minres_residual"]),
    }


def space_group_times(ed, cfg, steps, warmup):
    """The same J1-J2 Hamiltonian on the k = (0,0) A1 sector of the space group
    (translations x C4v), without and with the even spin flip: dimension,
    reduction time, FP32/FP64 matvec time and mixed-precision time to E0."""
    from edhip import _lib

    Lx, Ly = cfg["Lx"], cfg["Ly"]
    if Lx != Ly or tuple(cfg["k"]) != (0, 0):
        return None
    n = Lx * Ly
    hs, _ = ed.spin_half_system(n)
    j1, j2 = ed.square_bonds(Lx, Ly)
    terms = [ed.heisenberg_bond(hs, i, j, 1.0) for i, j in j1] + [ed.heisenberg_bond(hs, i, j, cfg["J2"]) for i, j in j2]
    H = ed.simplify(sum(terms, ed.NullOperator()))
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    out = {}
    base = ed.SquareSpaceIrrep(Lx, (0, 0), "A1")
    flip = ed.make_product_irrep([(ed.GlobalBitFlip(False), 1.0), (ed.GlobalBitFlip(True), 1.0)], base)
    for name, ops, order in (("A1", base, 8 * n), ("A1_even_flip", flip, 16 * n)):
        t0 = time.perf_counter()
        r = ed.symmetry_reduce(hsr, ops)
        t_red = time.perf_counter() - t0
        opr = ed.represent(r, H)
        ms32, _ = time_matvecs(ed, opr, _lib.F32, steps, warmup)
        ms64, _ = time_matvecs(ed, opr, _lib.F64, steps, warmup)
        # eigs assembles the rows of these sectors for the solve (entries of a
        # few GB); the matrix-free solve beside it
        res = ed.eigs(opr, 1, mode="mixed", tol=1e-10, max_basis=100, max_restarts=20, criterion="eigenvalue")
        mf = ed.eigs(opr, 1, mode="mixed", tol=1e-10, max_basis=100, max_restarts=20, criterion="eigenvalue",
                     assemble="off")
        opr.set_strategy("assembled")
        n_ent, n_bytes, t_build = opr.assembled_info()
        a32 = kernel_ms(ed, opr, _lib.F32, steps)
        a64 = kernel_ms(ed, opr, _lib.F64, steps)
        opr.set_strategy()
        out[name] = {"group_order": order, "dim": r.dimension, "reduce_seconds": t_red, "matvec_ms_f32": ms32,
                     "matvec_ms_f64": ms64, "E0": float(res.values[0]), "time_to_E0_mixed_s": res.t_total,
                     "assembled": res.assembled, "matvecs": res.matvecs, "matvecs_fp64": res.matvecs_fp64,
                     "time_to_E0_mixed_matrix_free_s": mf.t_total, "E0_matrix_free": float(mf.values[0]),
                     "assembled_entries": n_ent, "assembled_bytes": n_bytes, "assembled_build_s": t_build,
                     "assembled_kernel_ms_f32": a32, "assembled_kernel_ms_f64": a64}
        ed.release_workspace()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="6x6", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-e0", action="store_true",
                    help="skip the time-to-E0 runs (about 40 s; also those on the assembled rows)")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-shares", action="store_true", help="skip the eight row shares of the FP32 product")
    ap.add_argument("--no-assembled", action="store_true",
                    help="skip the assembled-row strategy (77 GB of entries at 6x6; about 20 s)")
    ap.add_argument("--tol-fp32", type=float, default=0.0,
                    help="mixed Lanczos: eigenvalue tolerance of the FP32 phase (0: max(tol, 1e-10))")
    ap.add_argument("--no-space-group", action="store_true",
                    help="skip the space-group (translations x C4v, and x the even spin flip) A1 sectors (about 15 s)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one plain 1 GiB device copy (elementwise kernel) for HBM-counter calibration and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        x = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        torch.mul(x, 1.0, out=y)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes_read": x.numel() * 4, "bytes_written": y.numel() * 4}))
        return
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU: there is no multi-GPU path (asked for {args.gpus}); "
            "the representatives would shard by contiguous ranges, with an all-gather of x before each product",
            file=sys.stderr,
        )
        sys.exit(2)

    import __graft_entry__

    ed = __graft_entry__.load_package()
    from edhip import _lib

    cfg = CONFIGS[args.config]
    hs, H, hsr, rhsr, opr, t_reduce = build_workload(ed, cfg)
    D = opr.dimension
    ni = ctypes.c_int64()
    opr.ctx.check(_lib.lib().edhip_op_count_images(opr._handle, ctypes.byref(ni)))
    n_img = int(ni.value)
    elems = D + n_img

    ms32, cs32 = time_matvecs(ed, opr, _lib.F32, args.steps, args.warmup)
    ms64, cs64 = time_matvecs(ed, opr, _lib.F64, args.steps, args.warmup)
    kms32 = kernel_ms(ed, opr, _lib.F32, args.steps)

    def b_alg(sT):
        return D * (8 + 2 * sT) + n_img * (8 + sT)

    # the dominant kernel alone (the product also spreads x before it)
    achieved = b_alg(4) / (kms32 * 1e-3) / 1e9
    traffic = None
    traffic_all = {}
    tfile = os.path.join(ROOT, "profiles", f"traffic_matvec_{args.config}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic_all = json.load(f)
        traffic = traffic_all.get("hbm_bytes_per_launch_f32")

    line = {
        "metric": "matvec Gelem/s + time-to-E0 @1e-10 rel, FP32-mixed vs FP64 GPU; grad time",
        "value": elems / (ms32 * 1e-3) / 1e9,
        "unit": "Gelem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 vectors (FP32 products), u64 state words, i64 indices",
        "data": "synthetic: deterministic J1-J2 Hamiltonian, seeded Gaussian input vector",
        "config": {
            "workload": f"J1-J2 Heisenberg {cfg['Lx']}x{cfg['Ly']} torus, J2={cfg['J2']}, Sz=0, "
            f"k={cfg['k']} translation sector, matrix-free matvec",
            "dim": D,
            "offdiag_images_per_matvec": n_img,
            "elements_per_matvec": elems,
            "parallelism": "single GPU",
        },
        "fp64_matvec": {"ms_per_step": ms64, "Gelem_per_s": elems / (ms64 * 1e-3) / 1e9},
        "reduce_seconds": t_reduce,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_matvec<float> (term list, 6x6 row-pair canonicaliser, x / sqrt(N) spread over word positions)",
            "kernel_ms": kms32,
            # measured HBM bytes (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated) over this run's kernel time
            "traffic_GBps": traffic / (kms32 * 1e-3) / 1e9 if traffic else None,
            "traffic_frac": traffic / (kms32 * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
        },
        "checksum_f32": cs32,
        "checksum_f64": cs64,
    }
    if not args.no_shares:
        line["row_shares"] = share_times(ed, opr)
    if not args.no_e0:
        line["time_to_E0"] = lanczos_times(ed, opr, args.tol_fp32)
        # the same runs with every step orthogonalised against the whole basis
        full = lanczos_times(ed, opr, args.tol_fp32, reorth="adaptive")
        line["time_to_E0_full_reorth"] = {
            m: {k: full[m][k] for k in ("E0", "seconds", "matvecs", "ortho_kernel_seconds", "full_reorth_steps")}
            for m in ("fp64", "mixed")
        }
        line["time_to_E0_full_reorth"]["speedup_mixed_vs_fp64"] = full["speedup_mixed_vs_fp64"]
    if not args.no_assembled:
        asm = assembled_times(ed, opr, elems, args.steps, args.warmup, e0=not args.no_e0)
        for sfx in ("f32", "f64"):
            tb = traffic_all.get(f"assembled_hbm_bytes_per_launch_{sfx}")
            if tb:
                asm[sfx]["hbm_bytes"] = tb
                asm[sfx]["hbm_GBps"] = tb / (asm[sfx]["kernel_ms"] * 1e-3) / 1e9
                asm[sfx]["hbm_frac"] = asm[sfx]["hbm_GBps"] / HBM_PEAK_GBS
        line["assembled_rows"] = asm
    if not args.no_grad:
        line["grad"] = grad_times(ed)
    if not args.no_space_group:
        line["space_group"] = space_group_times(ed, cfg, args.steps, args.warmup)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ed)
        line["small_configs_gpu"] = small_config_times(ed)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
