"""Brute-force dense checker (n <= 10) - TEST INFRASTRUCTURE ONLY.

Builds the full 2^n x 2^n operator of every gate as a Kronecker product of
2x2 blocks (or a permuted 4x4 block for two-qubit gates) and multiplies the
state vector.  This construction shares no code with statevector.py
(tensordot contraction), so agreement between the two pins the bit order
(qubit q = most-significant-first axis q, engine.py:1049-1057) and the
two-qubit matrix convention (basis |q0 q1>, gates.py:82-90) independently.
"""
from __future__ import annotations

from functools import reduce
from typing import Sequence, Tuple

import numpy as np

C128 = np.complex128
_P = {
    "I": np.eye(2, dtype=C128),
    "X": np.array([[0, 1], [1, 0]], dtype=C128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=C128),
    "Z": np.array([[1, 0], [0, -1]], dtype=C128),
}


def op_1q(U: np.ndarray, q: int, n: int) -> np.ndarray:
    mats = [np.eye(2, dtype=C128)] * n
    mats = list(mats)
    mats[q] = np.asarray(U, dtype=C128)
    return reduce(np.kron, mats)


def op_2q(U4: np.ndarray, q0: int, q1: int, n: int) -> np.ndarray:
    """Full operator of a 4x4 gate on (q0, q1), q0 the most significant bit of
    the 4x4 basis: element-wise definition over basis states."""
    dim = 1 << n
    U4 = np.asarray(U4, dtype=C128)
    out = np.zeros((dim, dim), dtype=C128)
    b0, b1 = n - 1 - q0, n - 1 - q1
    rest_mask = ~((1 << b0) | (1 << b1))
    for col in range(dim):
        c = (((col >> b0) & 1) << 1) | ((col >> b1) & 1)
        base = col & rest_mask
        for r in range(4):
            row = base | (((r >> 1) & 1) << b0) | ((r & 1) << b1)
            out[row, col] += U4[r, c]
    return out


def pauli_string(paulis: Sequence[Tuple[str, int]], n: int) -> np.ndarray:
    mats = [_P["I"]] * n
    mats = list(mats)
    for P, q in paulis:
        mats[int(q)] = _P[P]
    return reduce(np.kron, mats)


def z_expect(psi: np.ndarray, q: int, n: int) -> float:
    Zq = pauli_string([("Z", q)], n)
    return float(np.real(np.conj(psi) @ (Zq @ psi)))


def pauli_sum_dense(terms, n: int) -> np.ndarray:
    H = np.zeros((1 << n, 1 << n), dtype=C128)
    for c, fac in terms:
        H = H + float(c) * pauli_string(fac, n)
    return H
