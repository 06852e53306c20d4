import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

ED = __graft_entry__.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP kernels through the C ABI)")


def _gpu_present():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_present():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def ed():
    return ED


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    o.lib()
    return o


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "reference_known_answers.json")) as f:
        return json.load(f)


# ---- helpers shared by the tests -------------------------------------------


def heisenberg_terms_pauli(bonds, weights=None):
    """Hand-written PureOperator terms of sum_b w_b sigma_i . sigma_j (Pauli
    units): per bond four diagonal projectors (+1, -1, -1, +1) and two flips
    (amplitude 2), independent of the product's operator algebra."""
    mask, row, col, amp = [], [], [], []
    for k, (i, j) in enumerate(bonds):
        w = 1.0 if weights is None else weights[k]
        bi, bj = 1 << i, 1 << j
        m = bi | bj
        for r, c, a in ((0, 0, 1.0), (bi, bi, -1.0), (bj, bj, -1.0), (m, m, 1.0), (bi, bj, 2.0), (bj, bi, 2.0)):
            mask.append(m)
            row.append(r)
            col.append(c)
            amp.append(w * a)
    return (np.array(mask, np.uint64), np.array(row, np.uint64), np.array(col, np.uint64), np.array(amp, complex))


def dense_from_terms(n_bits, words, terms):
    """<b|O|b'> over the given words by the PureOperator row rule
    (operator_iterator.jl:34-38), straight Python loops for small cases."""
    mask, row, col, amp = terms
    pos = {int(w): i for i, w in enumerate(words)}
    M = np.zeros((len(words), len(words)), complex)
    for i, b in enumerate(words):
        b = int(b)
        for t in range(len(mask)):
            m = int(mask[t])
            if (b & m) == int(row[t]):
                bc = (b & ~m) | int(col[t])
                if bc in pos:
                    M[i, pos[bc]] += amp[t]
    return M


def chain_bonds(L):
    return [(i, (i + 1) % L) for i in range(L)]


def translation_group_perms(Lx, Ly=1):
    """0-based image sites of T(a, b) for every (a, b), identity first."""
    out = []
    for b in range(Ly):
        for a in range(Lx):
            out.append([((i % Lx + a) % Lx) + Lx * ((i // Lx + b) % Ly) for i in range(Lx * Ly)])
    return np.array(out, np.int32)


def translation_chars(Lx, Ly, kx, ky):
    return np.array(
        [np.exp(-2j * np.pi * (kx * a / Lx + ky * b / Ly)) for b in range(Ly) for a in range(Lx)], complex
    )
