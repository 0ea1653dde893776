"""Helpers shared by the GPU parity tests (no test functions here)."""
from __future__ import annotations

import numpy as np


def rand_state(rng, n):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def to_gpu(v):
    import torch

    return torch.tensor(np.asarray(v, dtype=np.complex64)).cuda().contiguous()


def to_np(t):
    return t.detach().cpu().numpy().astype(np.complex128)


def rand_unitary(rng, d):
    q, r = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


ONE = ["h", "x", "s", "sdg", "rx", "ry", "rz", "y", "z", "t", "tdg"]
TWO = ["cx", "cz", "iswap", "swap", "rxx", "ryy", "rzz", "cry", "cy"]


def random_circuit_ops(rng, n, depth, p_two=0.4, allow=None):
    ops = []
    for _ in range(depth):
        if n >= 2 and rng.random() < p_two:
            a, b = (int(v) for v in rng.choice(n, 2, replace=False))
            name = TWO[rng.integers(len(TWO))]
            if allow is not None and name not in allow:
                name = "cx"
            if name in ("rxx", "ryy", "rzz", "cry"):
                ops.append((name, a, b, float(rng.uniform(0, 2 * np.pi))))
            else:
                ops.append((name, a, b))
        else:
            name = ONE[rng.integers(len(ONE))]
            if allow is not None and name not in allow:
                name = "h"
            q = int(rng.integers(n))
            if name in ("rx", "ry", "rz"):
                ops.append((name, q, float(rng.uniform(0, 2 * np.pi))))
            else:
                ops.append((name, q))
    return ops


def random_pauli_terms(rng, n, k, maxw=4):
    terms = []
    for _ in range(k):
        w = int(rng.integers(1, min(n, maxw) + 1))
        qs = rng.choice(n, w, replace=False)
        terms.append((float(rng.normal()), [("XYZ"[int(rng.integers(3))], int(q)) for q in qs]))
    return terms


# ---- arbitrary-matrix circuits (bit-level ops built directly) ----------------
def ops_from_items(n, items):
    """tqsv_op array from items ("m1", q, U2x2, cmask, cval) / ("m2", q0, q1, U4x4)."""
    from tyxonq_amd.engine import _Lowerer

    L = _Lowerer(n, False)
    for it in items:
        if it[0] == "m1":
            _, q, m, cm, cv = it
            L.mat1(q, m, cmask=cm, cval=cv)
        else:
            _, q0, q1, m = it
            L.mat2(q0, q1, m)
    L.flush()
    return L.segments[0][1]


def dense_apply_items(n, items, v):
    """The same items applied with the oracle's contractions (complex128)."""
    from oracle import statevector as O

    out = np.asarray(v, dtype=complex).copy()
    for it in items:
        if it[0] == "m1":
            _, q, m, cm, cv = it
            moved = O.apply_1q(out, m, q, n)
            if cm == 0:
                out = moved
            else:
                sel = (np.arange(1 << n) & cm) == cv
                out = np.where(sel, moved, out)
        else:
            _, q0, q1, m = it
            out = O.apply_2q(out, m, q0, q1, n)
    return out


def phase_variants(rng, u):
    """u times random global factors and the +-1/+-i one-sided phases that the
    planner's rotation forms and sign frames treat specially."""
    ph = [1, -1, 1j, -1j, np.exp(1j * rng.uniform(0, 2 * np.pi))]
    a, b = ph[int(rng.integers(5))], ph[int(rng.integers(5))]
    g = rng.uniform(0.5, 2.0) * np.exp(1j * rng.uniform(0, 2 * np.pi)) if rng.random() < 0.3 else 1
    return g * np.diag([1, a]) @ np.asarray(u) @ np.diag([1, b])


def random_matrix_items(rng, n, count):
    """Arbitrary 1-qubit unitaries (incl. scaled ones and the H / X / RX / RY
    special forms), diagonal gates on any two bits (sign-only and general
    phases), controlled and non-unitary 2x2 matrices and dense 4x4 unitaries."""
    items = []
    for _ in range(count):
        r = rng.random()
        q = int(rng.integers(n))
        other = int(rng.integers(n - 1))
        other = other if other < q else other + 1
        if r < 0.45:
            u = rand_unitary(rng, 2)
            if rng.random() < 0.3:
                c, s = np.cos(rng.uniform(0, 3)), np.sin(rng.uniform(0, 3))
                u = [np.array([[c, -s], [s, c]]), np.array([[c, -1j * s], [-1j * s, c]]),
                     np.array([[0, 1], [1, 0]]),
                     np.array([[1, 1], [1, -1]]) / np.sqrt(2)][int(rng.integers(4))]
            items.append(("m1", q, phase_variants(rng, u), 0, 0))
        elif r < 0.75:
            d = [1, 1, 1, -1] if rng.random() < 0.5 else np.exp(1j * rng.uniform(0, 2 * np.pi, 4))
            items.append(("m2", q, other, np.diag(d)))
        elif r < 0.85:
            cb = 1 << (n - 1 - other)
            items.append(("m1", q, rand_unitary(rng, 2), cb, cb if rng.random() < 0.7 else 0))
        elif r < 0.92:
            items.append(("m1", q, rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)), 0, 0))
        else:
            items.append(("m2", q, other, rand_unitary(rng, 4)))
    return items
