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
