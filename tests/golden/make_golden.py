"""Writes tests/golden/reference_known_answers.json.

Two kinds of data, both independent of the oracle and of the product:

1. Known answers transcribed from the reference's own test files (data only:
   inputs and expected outputs, each with its test/file:line).
2. Values fixed by mathematics, computed here by brute force with numpy on
   dense matrices built from Kronecker products of Pauli matrices, and by
   character sums (Burnside): sector dimensions and lowest eigenvalues.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import cmath
import json
import math
import os
from itertools import product

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------------------
# 1. transcribed known answers
# ---------------------------------------------------------------------------
def known_answers():
    return {
        "hsr_n4_sz0_basis": {
            "source": "test/test_hilbert_space_representation.jl:13, test/test_reduced_representation.jl:97",
            "n_sites": 4,
            "sz": 0,
            "basis": [0b0011, 0b0101, 0b0110, 0b1001, 0b1010, 0b1100],
        },
        "tj_sectors": {
            "source": "test/test_hilbert_space_representation.jl:73-128",
            "sites": [
                [["Em", [0, 0]], ["Up", [1, 1]], ["Dn", [1, -1]]],
                [["Up", [1, 1]], ["Dn", [1, -1]]],
                [["Up", [1, 1]], ["Dn", [1, -1]]],
            ],
            "sectors": [[2, -2], [2, 0], [2, 2], [3, -3], [3, -1], [3, 1], [3, 3]],
            "full_basis": [0b0000, 0b0001, 0b0010, 0b0100, 0b0101, 0b0110, 0b1000, 0b1001, 0b1010, 0b1100, 0b1101, 0b1110],
            "sector_basis": {
                "2,-2": [0b1100],
                "2,0": [0b0100, 0b1000],
                "2,2": [0b0000],
                "3,-3": [0b1110],
                "3,-1": [0b0110, 0b1010, 0b1101],
                "3,1": [0b0010, 0b0101, 0b1001],
                "3,3": [0b0001],
            },
            "two_sector_basis": {"sectors": [[3, 3], [3, 1]], "basis": [0b0001, 0b0010, 0b0101, 0b1001]},
        },
        "translation_n4": {
            "source": "test/test_symmetry_reduce.jl:33-48",
            "n_sites": 4,
            "reduced_basis_by_irrep_index": {"1": [0b0011, 0b0101], "2": [0b0011], "3": [0b0011, 0b0101], "4": [0b0011]},
        },
        "inversion_n4": {
            "source": "test/test_symmetry_reduce.jl:102-107 (point group '-1' on a 4-site chain: site map [1,4,3,2])",
            "map": [1, 4, 3, 2],
            "reduced_basis": {"even": [0b0011, 0b0101, 0b0110, 0b1010], "odd": [0b0011, 0b0110]},
        },
        "space_group_n4": {
            "source": "test/test_symmetry_reduce.jl:128-141 (translations x inversion, k=0; inversion even / odd)",
            "reduced_basis": {"k0_even": [0b0011, 0b0101], "k0_odd": []},
        },
        "phase_convention_n7": {
            "source": "test/test_symmetry_reduce.jl:164-214",
            "n_sites": 7,
            "plus5": {
                "basis": [0b0000001, 0b0000010, 0b0000100, 0b0001000, 0b0010000, 0b0100000, 0b1000000],
                "reduced": [0b0000001],
                "unreduce_of_one": "cis(2 pi i / 7) / sqrt(7), i = 0..6",
            },
            "minus5": {
                "basis": [0b0111111, 0b1011111, 0b1101111, 0b1110111, 0b1111011, 0b1111101, 0b1111110],
                "reduced": [0b0111111],
                "unreduce_of_one": "cis(-2 pi i / 7) / sqrt(7), i = 0..6",
            },
            "irrep_index": 2,
        },
        "symmetry_apply": {
            "source": "test/test_symmetry_apply.jl:20-41",
            "transop": [2, 3, 4, 1],
            "invop": [1, 4, 3, 2],
            "bits": [[0b0001, 0b0010], [0b0101, 0b1010]],
            "pop1": [0b1101, 0b0101, 0b1100, 2.0],
            "pop1_trans": [0b1011, 0b1010, 0b1001, 2.0],
            "pop1_inv": [0b0111, 0b0101, 0b0110, 2.0],
        },
        "iterator_invalid": {
            "source": "test/test_operator_representation.jl:123-132",
            "terms": [[0b10, 0b00, 0b10, 2.0], [0b10, 0b10, 0b00, 3.0]],
            "basis": [0b00, 0b01],
            "row_iterator": {"0": [[-1, 2.0]], "1": [[-1, 2.0]]},
            "column_iterator": {"0": [[-1, 3.0]], "1": [[-1, 3.0]]},
        },
        "heisenberg_n4_matrix_pauli": {
            "source": "test/test_operator_application.jl:142-147 (commented out in the reference; re-derived below)",
            "matrix": [
                [0, 2, 0, 0, 2, 0],
                [2, -4, 2, 2, 0, 2],
                [0, 2, 0, 0, 2, 0],
                [0, 2, 0, 0, 2, 0],
                [2, 0, 2, 2, -4, 2],
                [0, 2, 0, 0, 2, 0],
            ],
        },
        "reduced_2x2_n4_k0": {
            "source": "test/test_reduced_representation.jl:159-165",
            "psis": [[1, 0, 1, 1, 0, 1], [0, 1, 0, 0, 1, 0]],
        },
    }


# ---------------------------------------------------------------------------
# 2. brute force (dense Kronecker products) and closed forms
# ---------------------------------------------------------------------------
SX = np.array([[0, 1], [1, 0]], dtype=complex)
SY = np.array([[0, -1j], [1j, 0]], dtype=complex)
SZ = np.array([[1, 0], [0, -1]], dtype=complex)
I2 = np.eye(2, dtype=complex)


def site_op(n, i, m):
    """Operator m on site i (0-based).  Word index = sum_i bit_i 2^i with
    local state 0 (Up) at bit 0, so site 0 is the rightmost Kronecker factor
    and local index 0 = Up (the Pauli matrices above act on (Up, Dn))."""
    out = np.array([[1.0 + 0j]])
    for j in reversed(range(n)):
        out = np.kron(out, m if j == i else I2)
    return out


def heisenberg_dense(n, bonds, weights=None, pauli_units=False):
    dim = 2 ** n
    H = np.zeros((dim, dim), dtype=complex)
    for k, (i, j) in enumerate(bonds):
        w = 1.0 if weights is None else weights[k]
        for m in (SX, SY, SZ):
            H += w * site_op(n, i, m) @ site_op(n, j, m)
    return H if pauli_units else 0.25 * H


def sz0_block(H, n):
    idx = [b for b in range(2 ** n) if bin(b).count("1") == n // 2]
    return H[np.ix_(idx, idx)], idx


def heisenberg_sparse_sz(n, bonds, weights=None):
    """Sparse Sz=0 Heisenberg (S units) built with bit tricks (numpy)."""
    import scipy.sparse as sp

    words = np.array([b for b in range(2 ** n) if bin(b).count("1") == n // 2], dtype=np.int64)
    pos = {int(b): i for i, b in enumerate(words)}
    rows, cols, vals = [], [], []
    diag = np.zeros(words.size)
    for k, (i, j) in enumerate(bonds):
        w = 1.0 if weights is None else weights[k]
        bi = (words >> i) & 1
        bj = (words >> j) & 1
        diag += np.where(bi == bj, 0.25 * w, -0.25 * w)
        flip = bi != bj
        src = np.nonzero(flip)[0]
        tgt = words[src] ^ ((1 << i) | (1 << j))
        rows.extend(src.tolist())
        cols.extend(pos[int(t)] for t in tgt)
        vals.extend([0.5 * w] * src.size)
    H = sp.csr_matrix((vals, (rows, cols)), shape=(words.size, words.size))
    return H + sp.diags(diag)


def translation_orbit_dims(Lx, Ly, kx, ky, n_up):
    """dim of the momentum sector by the character formula
    D_k = (1/|G|) sum_g conj(chi_k(g)) * #{Sz words fixed by g}."""
    n = Lx * Ly
    G = Lx * Ly
    tot = 0j
    for b_ in range(Ly):
        for a in range(Lx):
            # cycles of the site permutation T(a, b)
            seen = [False] * n
            cycles = []
            for s in range(n):
                if seen[s]:
                    continue
                c = 0
                t = s
                while not seen[t]:
                    seen[t] = True
                    c += 1
                    x, y = t % Lx, t // Lx
                    t = ((x + a) % Lx) + Lx * ((y + b_) % Ly)
                cycles.append(c)
            # words fixed by g with n_up set bits: choose whole cycles
            ways = [0] * (n + 1)
            ways[0] = 1
            for c in cycles:
                for m in range(n, c - 1, -1):
                    ways[m] += ways[m - c]
            chi = cmath.exp(-2j * math.pi * (kx * a / Lx + ky * b_ / Ly))
            tot += chi.conjugate() * ways[n_up]
    v = tot / G
    return int(round(v.real))


def main():
    data = {"known_answers": known_answers()}
    # brute-force checks of the transcribed n=4 matrix (Pauli units)
    n = 4
    bonds = [(i, (i + 1) % n) for i in range(n)]
    H4 = heisenberg_dense(n, bonds, pauli_units=True)
    blk, _ = sz0_block(H4, n)
    assert np.allclose(blk.real, data["known_answers"]["heisenberg_n4_matrix_pauli"]["matrix"])
    psis = [np.array(p, float) / np.linalg.norm(p) for p in data["known_answers"]["reduced_2x2_n4_k0"]["psis"]]
    H2 = np.array([[psis[i] @ blk.real @ psis[j] for j in range(2)] for i in range(2)])
    data["known_answers"]["reduced_2x2_n4_k0"]["matrix"] = H2.tolist()
    # sigma^+ on site 2 (1-based) of 4: kron(s0, s0, s+, s0) (test_operator_representation.jl:135-147)
    sp_ = np.array([[0, 1], [0, 0]])
    s0 = np.eye(2)
    data["known_answers"]["sigma_plus_site2_n4"] = {
        "source": "test/test_operator_representation.jl:134-147",
        "dense": np.kron(np.kron(np.kron(s0, s0), sp_), s0).astype(int).tolist(),
    }

    # lowest eigenvalues (S.S units)
    import scipy.sparse.linalg as sla

    ev = {}
    for L in (8, 10, 12, 16):
        Hs = heisenberg_sparse_sz(L, [(i, (i + 1) % L) for i in range(L)])
        w = np.sort(sla.eigsh(Hs, k=4, which="SA", tol=1e-13, return_eigenvectors=False))
        ev[f"chain{L}"] = w.tolist()
    # 4x4 square
    Lx = Ly = 4
    s = lambda x, y: (x % Lx) + Lx * (y % Ly)
    b44 = [(s(x, y), s(x + 1, y)) for y in range(Ly) for x in range(Lx)] + [
        (s(x, y), s(x, y + 1)) for y in range(Ly) for x in range(Lx)
    ]
    Hs = heisenberg_sparse_sz(16, b44)
    ev["square4x4"] = np.sort(sla.eigsh(Hs, k=1, which="SA", tol=1e-13, return_eigenvectors=False)).tolist()
    data["eigenvalues_brute_force"] = {
        "source": "scipy eigsh on sparse Sz=0 matrices built with bit operations (S.S units)",
        **ev,
    }
    # sector dimensions by character sums
    dims = {
        "chain16_sz0": math.comb(16, 8),
        "chain24_k3": translation_orbit_dims(24, 1, 3, 0, 12),
        "torus4x6_pi0": translation_orbit_dims(4, 6, 2, 0, 12),
        "torus6x4_pi0": translation_orbit_dims(6, 4, 3, 0, 12),
        "torus6x6_k0": translation_orbit_dims(6, 6, 0, 0, 18),
        "torus4x6_sz0": math.comb(24, 12),
        "chain12_by_k": [translation_orbit_dims(12, 1, k, 0, 6) for k in range(12)],
        "torus4x4_by_k": [[translation_orbit_dims(4, 4, kx, ky, 8) for ky in range(4)] for kx in range(4)],
    }
    data["sector_dimensions"] = {"source": "character sums over the translation group (Burnside)", **dims}
    path = os.path.join(HERE, "reference_known_answers.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
