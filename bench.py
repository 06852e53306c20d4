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
spread of x that precedes it excluded) against the 8 TB/s HBM peak.  ``cpu_baseline`` times the oracle's
apply_serial! (the reference algorithm: parent binary search + mapping
arrays, one core) on the largest J1-J2 instance that algorithm can build,
the 4x6 lattice at k = (pi, 0).

``grad`` times the backward pass on the 4x6 random-bond XXZ model.

``time_to_E0`` times the FP32-mixed and the FP64 Lanczos to E0 on the
workload's operator.

Options: --steps K --warmup W --gpus 1 [--config 4x6|6x6] [--no-cpu]
[--no-grad] [--no-e0] [--space-group].
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


def cpu_baseline(ed):
    """Oracle apply_serial! on the 4x6 J1-J2 (pi,0) sector, one core."""
    import oracle

    cfg = CONFIGS["4x6"]
    Lx, Ly = cfg["Lx"], cfg["Ly"]
    n = Lx * Ly
    hs, _ = ed.spin_half_system(n)
    j1, j2 = ed.square_bonds(Lx, Ly)
    terms = [ed.heisenberg_bond(hs, i, j, 1.0) for i, j in j1] + [ed.heisenberg_bond(hs, i, j, cfg["J2"]) for i, j in j2]
    H = ed.simplify(sum(terms, ed.NullOperator()))
    irrep = ed.TranslationIrrep((Lx, Ly), cfg["k"])
    parent = oracle.sector_basis(n, 0)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in irrep.symops_and_amplitudes()])
    chars = [a for _, a in irrep.symops_and_amplitudes()]
    t0 = time.perf_counter()
    red = oracle.symmetry_reduce(parent, n, perms, chars)
    t_reduce = time.perf_counter() - t0
    o = oracle.Opr(red, ed.terms_arrays(H))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(o.dim) + 0j
    # elements per matvec: diagonal + off-diagonal images generated
    n_off = 0
    mask, row, col = o.mask, o.row, o.col
    off = row != col
    for b in red.basis_list:
        n_off += int(np.count_nonzero((np.uint64(b) & mask[off]) == row[off]))
    reps = 0
    t0 = time.perf_counter()
    while True:
        o.matvec(x)
        reps += 1
        if time.perf_counter() - t0 > 10.0 or reps >= 50:
            break
    dt = (time.perf_counter() - t0) / reps
    return {
        "value": (o.dim + n_off) / dt / 1e9,
        "unit": "Gelem/s",
        "cores": 1,
        "kind": "port",
        "sample": (
            f"oracle apply_serial! (reference algorithm: parent-basis binary search + basis_mapping arrays) on the "
            f"4x6 J1-J2 k=(pi,0) sector, {o.dim} states, {reps} matvecs in {dt * reps:.1f}s on 1 core; the 6x6 "
            f"workload cannot be built by that algorithm (parent arrays over 9.08e9 states); "
            f"oracle symmetry_reduce took {t_reduce:.1f}s"
        ),
    }


def lanczos_times(ed, opr):
    """Time to E0 at 1e-10 relative (eigenvalue criterion r^2/gap) for the
    FP32-storage mixed Lanczos (with FP64 refinement) and the FP64 Lanczos."""
    out = {}
    for mode in ("fp64", "mixed"):  # the mixed run reuses the FP64 run's workspace
        r = ed.eigs(opr, 1, mode=mode, tol=1e-10, max_basis=100, max_restarts=20, criterion="eigenvalue")
        out[mode] = {
            "E0": float(r.values[0]),
            "seconds": r.t_total,
            "iterations": r.iterations,
            "matvecs": r.matvecs,
            "matvecs_fp64": r.matvecs_fp64,
            "matvec_seconds": r.t_matvec,
            "refine_seconds": r.t_refine,
            "reorth_second_passes": r.second_passes,
            "residual": float(r.residuals[0]),
            "converged": r.converged,
        }
    out["speedup_mixed_vs_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["rel_diff_E0"] = abs(out["mixed"]["E0"] - out["fp64"]["E0"]) / abs(out["fp64"]["E0"])
    return out


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
    t0 = time.perf_counter()
    for _ in range(reps):
        ex = model.expectations(psi, psi)
    t_bw = (time.perf_counter() - t0) / reps
    g = J.T @ ex.real
    y = torch.empty_like(psi)
    ed.mul_(y, H, psi)
    t0 = time.perf_counter()
    for _ in range(reps):
        ed.mul_(y, H, psi)
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
        res = ed.eigs(opr, 1, mode="mixed", tol=1e-10, max_basis=100, max_restarts=20, criterion="eigenvalue")
        out[name] = {"group_order": order, "dim": r.dimension, "reduce_seconds": t_red, "matvec_ms_f32": ms32,
                     "matvec_ms_f64": ms64, "E0": float(res.values[0]), "time_to_E0_mixed_s": res.t_total,
                     "matvecs": res.matvecs, "matvecs_fp64": res.matvecs_fp64}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="6x6", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-e0", action="store_true", help="skip the time-to-E0 runs (about 40 s)")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--space-group", action="store_true",
                    help="also time the space-group (translations x C4v, optionally x spin flip) A1 sectors")
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
            f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {args.gpus}); "
            "the sharding design is described in DESIGN.md",
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
    tfile = os.path.join(ROOT, "profiles", f"traffic_matvec_{args.config}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch_f32")

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
            "kernel": "k_matvec<float> (term list, 6x6 row-pair canonicaliser, x spread over word positions)",
            "kernel_ms": kms32,
        },
        "checksum_f32": cs32,
        "checksum_f64": cs64,
    }
    if not args.no_e0:
        line["time_to_E0"] = lanczos_times(ed, opr)
    if not args.no_grad:
        line["grad"] = grad_times(ed)
    if args.space_group:
        line["space_group"] = space_group_times(ed, cfg, args.steps, args.warmup)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ed)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
