"""Backward pass: Hellmann-Feynman energy gradients from the fused per-term
kernel, correlator gradients from the MINRES adjoint, and the autograd
Functions; checked against central finite differences (h = 1e-5) and an FP64
CPU reference gradient <psi|h_b|psi> (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _xxz(ed, Lx, Ly, seed):
    n = Lx * Ly
    hs, s = ed.spin_half_system(n)
    j1, _ = ed.square_bonds(Lx, Ly)
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    from edhip.grad import xxz_model

    model, fn = xxz_model(hs, hsr, j1)
    rng = np.random.default_rng(seed)
    B = len(j1)
    theta = np.concatenate([rng.uniform(0.5, 1.5, B), rng.uniform(0.8, 1.2, B)])
    return hs, s, hsr, j1, model, fn, theta


def _energy(ed, model, fn, theta, tol=1e-13):
    c, _ = fn(theta)
    return ed.eigs(model.hamiltonian(c), 1, tol=tol, max_basis=200).values[0]


def test_energy_gradient_vs_fd_small(ed):
    """4x3 torus (24 bonds, 48 parameters): every parameter by central FD."""
    from edhip.grad import energy_and_grad

    hs, s, hsr, bonds, model, fn, theta = _xxz(ed, 4, 3, 1)
    c, J = fn(theta)
    E0, g = energy_and_grad(model, c, J, tol=1e-13)
    h = 1e-5
    fd = np.zeros_like(theta)
    for p in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (_energy(ed, model, fn, tp) - _energy(ed, model, fn, tm)) / (2 * h)
    assert np.max(np.abs(g - fd)) < 1e-6
    # homogeneity in J at fixed Delta: sum_b J_b dE/dJ_b = E0
    B = len(bonds)
    assert abs(np.dot(theta[:B], g[:B]) - E0) < 1e-10 * abs(E0)


def test_energy_gradient_cfg5(ed, oracle):
    """4x6 random-bond XXZ, Sz = 0 (2,704,156 states), 96 parameters: the
    device gradient vs the FP64 CPU reference <psi|h_b|psi> (oracle apply of
    each bond operator) on a sample of bonds, vs central FD on a few
    parameters, and the homogeneity sum rule on all of them."""
    from edhip.grad import energy_and_grad, ground_state

    hs, s, hsr, bonds, model, fn, theta = _xxz(ed, 4, 6, 2024)
    c, J = fn(theta)
    H, r = ground_state(model, c, tol=1e-12)
    E0 = float(r.values[0])
    psi = r.vectors[0]
    ex = model.expectations(psi, psi).real
    g = J.T @ ex
    B = len(bonds)
    assert abs(np.dot(theta[:B], g[:B]) - E0) < 1e-9 * abs(E0)
    # CPU reference on a sample of bonds
    psi_h = psi.cpu().numpy().astype(np.complex128)
    parent = oracle.sector_basis(24, 0)
    for b in (0, 5, 7, 19, 26, 33, 40, 47):  # 16 of the 96 components
        K = oracle.Opr(parent, ed.terms_arrays(model.operators[b]))
        Z = oracle.Opr(parent, ed.terms_arrays(model.operators[B + b]))
        kb = np.vdot(psi_h, K.matvec(psi_h)).real
        zb = np.vdot(psi_h, Z.matvec(psi_h)).real
        dJ = kb + theta[B + b] * zb
        dD = theta[b] * zb
        assert abs(g[b] - dJ) < 1e-10
        assert abs(g[B + b] - dD) < 1e-10
    # central FD on two parameters
    h = 1e-5
    for p in (3, B + 11):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd = (_energy(ed, model, fn, tp, 1e-12) - _energy(ed, model, fn, tm, 1e-12)) / (2 * h)
        assert abs(fd - g[p]) < 1e-6


def test_correlator_gradient_minres_vs_fd(ed):
    """d<Sz_0 Sz_6>/dtheta through the MINRES adjoint vs central FD (4x3 torus)."""
    from edhip.grad import correlator_and_grad, ground_state

    hs, s, hsr, bonds, model, fn, theta = _xxz(ed, 4, 3, 3)
    O = ed.simplify(0.25 * s(0, "z") * s(6, "z"))
    c, J = fn(theta)
    val, g, info = correlator_and_grad(model, c, J, O, tol=1e-13, minres_tol=1e-13)
    assert info["minres_residual"] < 1e-10

    Oop = ed.represent(hsr, O)

    def corr(th):
        cc, _ = fn(th)
        _, rr = ground_state(model, cc, tol=1e-13)
        v = rr.vectors[0].cpu().numpy()
        return float(np.vdot(v, Oop @ v).real)

    assert abs(corr(theta) - val) < 1e-10
    h = 1e-5
    for p in (0, 5, 24 + 2, 24 + 17):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd = (corr(tp) - corr(tm)) / (2 * h)
        assert abs(fd - g[p]) < 1e-6, (p, fd, g[p])


def test_correlator_gradient_cfg5(ed, oracle):
    """4x6 random-bond XXZ (2,704,156 states): d<S^z_0 S^z_12>/dtheta through
    the MINRES adjoint against an FP64 CPU reference -2 Re <lambda|O_k|psi>
    (oracle apply of the bond operators on the device's lambda and psi) on
    sampled bonds, and against central FD (h = 1e-5) on two parameters."""
    from edhip.grad import correlator_and_grad, ground_state

    hs, s, hsr, bonds, model, fn, theta = _xxz(ed, 4, 6, 2024)
    B = len(bonds)
    O = ed.simplify(0.25 * s(0, "z") * s(12, "z"))
    c, J = fn(theta)
    val, g, info = correlator_and_grad(model, c, J, O, tol=1e-13, minres_tol=1e-12)
    assert info["minres_residual"] < 1e-10
    psi = info["psi"].cpu().numpy().astype(np.complex128)
    lam = info["lambda"].cpu().numpy().astype(np.complex128)
    parent = oracle.sector_basis(24, 0)
    for b in (2, 17, 30, 45):
        kb = np.vdot(lam, oracle.Opr(parent, ed.terms_arrays(model.operators[b])).matvec(psi)).real
        zb = np.vdot(lam, oracle.Opr(parent, ed.terms_arrays(model.operators[B + b])).matvec(psi)).real
        assert abs(g[b] - (-2.0 * (kb + theta[B + b] * zb))) < 1e-10
        assert abs(g[B + b] - (-2.0 * theta[b] * zb)) < 1e-10
    Oop = ed.represent(hsr, O)

    def corr(th):
        cc, _ = fn(th)
        _, rr = ground_state(model, cc, tol=1e-13)
        v = rr.vectors[0].cpu().numpy()
        return float(np.vdot(v, Oop @ v).real)

    assert abs(corr(theta) - val) < 1e-10
    h = 1e-5
    for p in (4, B + 23):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd = (corr(tp) - corr(tm)) / (2 * h)
        assert abs(fd - g[p]) < 1e-6, (p, fd, g[p])


def test_autograd_functions(ed):
    import torch
    from edhip.grad import _torch_function

    GroundStateEnergy, GroundStateCorrelator = _torch_function()
    hs, s, hsr, bonds, model, fn, theta = _xxz(ed, 4, 3, 5)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    E = GroundStateEnergy.apply(th, model, fn, {"tol": 1e-13})
    E.backward()
    from edhip.grad import energy_and_grad

    c, J = fn(theta)
    E0, g = energy_and_grad(model, c, J, tol=1e-13)
    assert abs(E.item() - E0) < 1e-11
    assert np.max(np.abs(th.grad.numpy() - g)) < 1e-10
    O = ed.simplify(0.25 * s(0, "z") * s(6, "z"))
    th2 = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    val = GroundStateCorrelator.apply(th2, model, fn, O, {"tol": 1e-13})
    (2.0 * val).backward()
    from edhip.grad import correlator_and_grad

    v, gc, _ = correlator_and_grad(model, c, J, O, tol=1e-13, minres_tol=1e-13)
    assert abs(val.item() - v) < 1e-10
    assert np.max(np.abs(th2.grad.numpy() - 2.0 * gc)) < 1e-8


def test_term_expectations_bitwise_repeatable(ed):
    """The fused per-term pass sums in a fixed order: two runs on the same
    vectors give identical bits (4x6 XXZ, 2,704,156 states, 288 terms)."""
    from edhip.grad import ground_state

    hs, s, hsr, bonds, model, fn, theta = _xxz(ed, 4, 6, 7)
    c, J = fn(theta)
    H, r = ground_state(model, c, tol=1e-10)
    psi = r.vectors[0]
    a = model.expectations(psi, psi)
    b = model.expectations(psi, psi)
    assert np.array_equal(a, b)
    import torch

    phi = torch.roll(psi, 1)
    x = model.expectations(phi, psi)
    y = model.expectations(phi, psi)
    assert np.array_equal(x, y)


@pytest.mark.parametrize("k,dtype", [((1, 0), "complex128"), ((1, 0), "complex64"), ((0, 0), "float64"),
                                     ((2, 2), "float32")])
@pytest.mark.parametrize("lookup", ["auto", "bitmap"])
def test_term_expectations_reduced_vs_oracle(ed, oracle, k, dtype, lookup):
    """The per-term pass on a reduced space (diagonal-term kernel and
    off-diagonal pass; psi of the images read through the spread of x or the
    bitmap rank index): 4x4 J1-J2 torus with a longitudinal and a transverse
    field, translation sector k, every term's <phi|T_t|psi> (unit amplitude)
    against the oracle's reduced representation of that term alone."""
    import torch

    Lx = Ly = 4
    n = Lx * Ly
    hs, s = ed.spin_half_system(n)
    j1, j2 = ed.square_bonds(Lx, Ly)
    H = ed.simplify(sum([ed.heisenberg_bond(hs, i, j, 1.0) for i, j in j1]
                        + [ed.heisenberg_bond(hs, i, j, 0.5) for i, j in j2]
                        + [0.3 * s(i, "z") + 0.15 * s(i, "x") for i in range(n)], ed.NullOperator()))
    irrep = ed.TranslationIrrep((Lx, Ly), k)
    rh = ed.symmetry_reduce(ed.represent(ed.HilbertSpaceSector(hs, 0)), irrep)
    perms = np.stack([ed.bit_permutation(hs, p) for p, _ in irrep.symops_and_amplitudes()])
    red = oracle.symmetry_reduce(oracle.sector_basis(n, 0), n, perms, [a for _, a in irrep.symops_and_amplitudes()])
    opr = ed.represent(rh, H)
    opr.set_strategy("terms", lookup)
    D = opr.dimension
    rng = np.random.default_rng(5)
    phi = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    psi = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    if not dtype.startswith("complex"):
        phi, psi = phi.real.copy(), psi.real.copy()
    tdt = getattr(torch, dtype)
    phi_d = torch.tensor(phi, dtype=tdt, device="cuda")
    psi_d = torch.tensor(psi, dtype=tdt, device="cuda")
    got = ed.term_expectations(opr, phi_d, psi_d)
    # the values the device saw (rounded to the storage type)
    phi_c = phi_d.cpu().numpy().astype(np.complex128)
    psi_c = psi_d.cpu().numpy().astype(np.complex128)
    mask, row, col, _ = ed.terms_arrays(H)
    tol = 1e-12 if dtype in ("complex128", "float64") else 1e-5
    scale = np.linalg.norm(phi_c) * np.linalg.norm(psi_c)
    assert got.size == mask.size
    for t in range(mask.size):
        o = oracle.Opr(red, (mask[t:t + 1], row[t:t + 1], col[t:t + 1], np.ones(1, complex)))
        ref = np.vdot(phi_c, o.matvec(psi_c))
        assert abs(got[t] - ref) <= tol * scale, (t, got[t], ref)


@pytest.mark.parametrize("dtype", ["complex128", "complex64", "float64", "float32"])
@pytest.mark.parametrize("space,lookup", [("sector", "auto"), ("sector", "search"), ("two_sectors", "auto"),
                                          ("list", "auto"), ("list", "search")])
def test_term_expectations_unreduced_vs_oracle(ed, oracle, dtype, space, lookup):
    """The per-term pass without a symmetry group (the lanes take the
    off-diagonal terms in lock-step): a 14-site chain with random XXZ bonds,
    next-nearest bonds and a longitudinal and transverse field (images that
    leave the space), on the Sz = 0 sector (rank arithmetic, with and without
    the spread of psi), on the quantum numbers {0, 2} (two popcounts) and on a random
    list of words (sorted search, bitmap rank index); every term's
    <phi|T_t|psi> (unit amplitude) against the oracle's representation of
    that term alone."""
    import torch

    n = 14
    hs, s = ed.spin_half_system(n)
    rng = np.random.default_rng(23)
    bonds = [(i, (i + 1) % n) for i in range(n)] + [(i, (i + 2) % n) for i in range(n)]
    H = ed.simplify(sum([ed.heisenberg_bond(hs, i, j, rng.uniform(0.5, 1.5), rng.uniform(0.8, 1.2)) for i, j in bonds]
                        + [0.3 * s(i, "z") + 0.15 * s(i, "x") for i in range(n)], ed.NullOperator()))
    if space == "sector":
        rep = ed.represent(ed.HilbertSpaceSector(hs, 0))
        basis = oracle.sector_basis(n, 0)
    elif space == "two_sectors":
        rep = ed.represent(ed.HilbertSpaceSector(hs, [0, 2]))
        basis = np.sort(np.concatenate([oracle.sector_basis(n, 0), oracle.sector_basis(n, 2)]))
    else:
        basis = np.sort(rng.choice(1 << n, 5000, replace=False).astype(np.uint64))
        rep = ed.represent(hs, basis)
    assert np.array_equal(rep.basis_list, basis)
    opr = ed.represent(rep, H)
    opr.set_strategy("terms", lookup)
    D = opr.dimension
    phi = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    psi = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    if not dtype.startswith("complex"):
        phi, psi = phi.real.copy(), psi.real.copy()
    tdt = getattr(torch, dtype)
    phi_d = torch.tensor(phi, dtype=tdt, device="cuda")
    psi_d = torch.tensor(psi, dtype=tdt, device="cuda")
    got = ed.term_expectations(opr, phi_d, psi_d)
    phi_c = phi_d.cpu().numpy().astype(np.complex128)
    psi_c = psi_d.cpu().numpy().astype(np.complex128)
    mask, row, col, _ = ed.terms_arrays(H)
    assert (row != col).sum() > 16  # more than one window of off-diagonal terms
    tol = 1e-12 if dtype in ("complex128", "float64") else 1e-5
    scale = np.linalg.norm(phi_c) * np.linalg.norm(psi_c)
    assert got.size == mask.size
    for t in range(mask.size):
        o = oracle.Opr(basis, (mask[t:t + 1], row[t:t + 1], col[t:t + 1], np.ones(1, complex)))
        ref = np.vdot(phi_c, o.matvec(psi_c))
        assert abs(got[t] - ref) <= tol * scale, (t, got[t], ref)


def test_couplings_on_reduced_space(ed):
    """On a translation sector, per-bond couplings are rejected (their reduced
    expectations would not be the physical ones); translation-invariant sums
    pass, and their Hellmann-Feynman gradient obeys E0 = sum_k c_k dE0/dc_k
    and matches the unreduced Sz sector's at the same couplings (the ring's
    ground state has k = 0)."""
    from edhip.grad import LinearCouplings, energy_and_grad, xxz_model

    L = 12
    hs, s = ed.spin_half_system(L)
    bonds = [(i, (i + 1) % L) for i in range(L)]
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    rh = ed.symmetry_reduce(hsr, ed.TranslationIrrep(L, 0))
    with pytest.raises(ed.ArgumentError):
        xxz_model(hs, rh, bonds)
    K = ed.simplify(sum((0.25 * (s(i, "x") * s(j, "x") + s(i, "y") * s(j, "y")) for i, j in bonds), ed.NullOperator()))
    Z = ed.simplify(sum((0.25 * s(i, "z") * s(j, "z") for i, j in bonds), ed.NullOperator()))
    c = np.array([1.0, 0.7])
    E_r, g_r = energy_and_grad(LinearCouplings(rh, [K, Z]), c, np.eye(2), tol=1e-13)
    E_u, g_u = energy_and_grad(LinearCouplings(hsr, [K, Z]), c, np.eye(2), tol=1e-13)
    assert abs(np.dot(c, g_r) - E_r) < 1e-11 * abs(E_r)
    assert abs(E_r - E_u) < 1e-11 * abs(E_u)
    assert np.max(np.abs(g_r - g_u)) < 1e-8


def test_correlator_gradient_on_reduced_space(ed):
    """A translation-invariant correlator on the k = 0 sector of a 12-site
    ring (O = (1/L) sum_i Sz_i Sz_{i+2}): its ground-state value and its
    MINRES-adjoint gradient with respect to the two couplings (J_xy, J_z)
    match the unreduced Sz sector's (the ground state lies at k = 0) and
    central finite differences (step 1e-5) on the reduced space."""
    from edhip.grad import LinearCouplings, correlator_and_grad, ground_state

    L = 12
    hs, s = ed.spin_half_system(L)
    bonds = [(i, (i + 1) % L) for i in range(L)]
    hsr = ed.represent(ed.HilbertSpaceSector(hs, 0))
    rh = ed.symmetry_reduce(hsr, ed.TranslationIrrep(L, 0))
    K = ed.simplify(sum((0.25 * (s(i, "x") * s(j, "x") + s(i, "y") * s(j, "y")) for i, j in bonds), ed.NullOperator()))
    Z = ed.simplify(sum((0.25 * s(i, "z") * s(j, "z") for i, j in bonds), ed.NullOperator()))
    O = ed.simplify(sum(((0.25 / L) * s(i, "z") * s((i + 2) % L, "z") for i in range(L)), ed.NullOperator()))
    c = np.array([1.0, 0.7])
    v_r, g_r, _ = correlator_and_grad(LinearCouplings(rh, [K, Z]), c, np.eye(2), O, tol=1e-13, minres_tol=1e-12)
    v_u, g_u, _ = correlator_and_grad(LinearCouplings(hsr, [K, Z]), c, np.eye(2), O, tol=1e-13, minres_tol=1e-12)
    assert abs(v_r - v_u) < 1e-10
    assert np.max(np.abs(g_r - g_u)) < 1e-7
    model = LinearCouplings(rh, [K, Z])
    Orep = ed.represent(rh, O)
    h = 1e-5

    def value(cc):
        H, r = ground_state(model, cc, tol=1e-13)
        psi = r.vectors[0].cpu().numpy()
        return float(np.vdot(psi, Orep @ psi).real)

    for k in range(2):
        e = np.zeros(2)
        e[k] = h
        fd = (value(c + e) - value(c - e)) / (2 * h)
        assert abs(fd - g_r[k]) < 1e-6, (k, fd, g_r[k])
