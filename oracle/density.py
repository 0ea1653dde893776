"""Density-matrix engine: NumPy complex128 restatement - TEST INFRASTRUCTURE
ONLY (see oracle/__init__.py).

Follows src/tyxonq/devices/simulators/density_matrix/engine.py:33-223 and the
kernels it calls (libs/quantum_library/kernels/density_matrix.py:9-110):
rho = |0..0><0..0|, U rho U^dag per gate (row axes first, qubit 0 the most
significant bit), sum_k K rho K^dag per channel, per-gate noise on the gate's
wires when use_noise is set (engine.py:185-215), projection with
renormalisation for project_z and reset (engine.py:216-223), <Z_q> from the
diagonal (density_matrix.py:66-73).  Channel Kraus sets follow
libs/quantum_library/noise.py:29-252.

How it is pinned (tests/test_density_oracle.py): pure states against the
statevector oracle (rho = |psi><psi|), trace preservation of every channel,
the closed forms <Z> = 1 - 4p/3 after depolarizing |0> and
<Z> = 1 - 2(1 - gamma) after amplitude damping of |1>, and the reference's
own known answer for the depolarizing Rx(pi) circuit.
"""
from __future__ import annotations

from typing import Any, Dict, List, Sequence

import numpy as np

from . import dense as D
from . import statevector as O

C128 = np.complex128
I2 = np.eye(2, dtype=C128)

# gates the reference's density-matrix run applies (engine.py:45-87); others are skipped
DM_GATES = {"h", "rz", "rx", "ry", "cx", "cz", "cry", "x", "s", "sdg"}


def depolarizing(p: float) -> List[np.ndarray]:  # noise.py:29-70
    s = np.sqrt(p / 3.0)
    return [np.sqrt(1 - p) * I2, s * O.X, s * O.Y, s * O.Z]


def amplitude_damping(g: float) -> List[np.ndarray]:  # noise.py:73-108
    return [np.array([[1, 0], [0, np.sqrt(1 - g)]], dtype=C128),
            np.array([[0, np.sqrt(g)], [0, 0]], dtype=C128)]


def phase_damping(lam: float) -> List[np.ndarray]:  # noise.py:111-147
    return [np.array([[1, 0], [0, np.sqrt(1 - lam)]], dtype=C128),
            np.array([[0, 0], [0, np.sqrt(lam)]], dtype=C128)]


def pauli_channel(px: float, py: float, pz: float) -> List[np.ndarray]:  # noise.py:150-205
    return [np.sqrt(1 - px - py - pz) * I2, np.sqrt(px) * O.X, np.sqrt(py) * O.Y,
            np.sqrt(pz) * O.Z]


def noise_kraus(noise: Dict[str, Any]) -> List[np.ndarray]:
    """The channel applied after each gate (engine.py:185-215)."""
    t = str(noise.get("type", "")).lower()
    if t == "depolarizing":
        return depolarizing(float(noise.get("p", 0.0)))
    if t == "amplitude_damping":
        return amplitude_damping(float(noise.get("gamma", noise.get("g", 0.0))))
    if t == "phase_damping":
        return phase_damping(float(noise.get("lambda", noise.get("l", 0.0))))
    if t == "pauli":
        return pauli_channel(float(noise.get("px", 0.0)), float(noise.get("py", 0.0)),
                             float(noise.get("pz", 0.0)))
    return []


def _full(op_on_qubits: np.ndarray, qubits: Sequence[int], n: int) -> np.ndarray:
    """Dense 2^n operator of a 1- or 2-qubit matrix on `qubits` (qubit 0 = MSB)."""
    if len(qubits) == 1:
        return D.op_1q(op_on_qubits, int(qubits[0]), n)
    return D.op_2q(op_on_qubits, int(qubits[0]), int(qubits[1]), n)


def apply_unitary(rho: np.ndarray, U: np.ndarray, qubits: Sequence[int], n: int) -> np.ndarray:
    M = _full(U, qubits, n)  # density_matrix.py:20-63: U rho U^dag
    return M @ rho @ M.conj().T


def apply_kraus(rho: np.ndarray, kops: Sequence[np.ndarray], q: int, n: int) -> np.ndarray:
    out = np.zeros_like(rho)  # density_matrix.py:76-110: sum_k K rho K^dag
    for K in kops:
        M = _full(np.asarray(K, dtype=C128), [q], n)
        out += M @ rho @ M.conj().T
    return out


def project(rho: np.ndarray, q: int, keep: int, n: int) -> np.ndarray:
    P = np.diag([1.0, 0.0] if keep == 0 else [0.0, 1.0]).astype(C128)  # engine.py:216-223
    r = apply_unitary(rho, P, [q], n)
    tr = np.trace(r)
    return r / tr if abs(tr) > 0 else r


def _gate(op) -> tuple:
    name = op[0]
    if name in ("h", "x", "s", "sdg", "y", "z", "t", "tdg"):
        m = {"h": O.gate_h(), "x": O.X, "s": np.diag([1, 1j]), "sdg": np.diag([1, -1j]),
             "y": O.Y, "z": O.Z, "t": np.diag([1, np.exp(1j * np.pi / 4)]),
             "tdg": np.diag([1, np.exp(-1j * np.pi / 4)])}[name]
        return np.asarray(m, dtype=C128), [int(op[1])]
    if name in ("rx", "ry", "rz"):
        f = {"rx": O.gate_rx, "ry": O.gate_ry, "rz": O.gate_rz}[name]
        return f(float(op[2])), [int(op[1])]
    if name in ("cx", "cz", "cy"):
        t = {"cx": O.X, "cz": O.Z, "cy": O.Y}[name]
        U = np.eye(4, dtype=C128)
        U[2:, 2:] = t
        return U, [int(op[1]), int(op[2])]
    if name == "cry":
        U = np.eye(4, dtype=C128)
        U[2:, 2:] = O.gate_ry(float(op[3]))
        return U, [int(op[1]), int(op[2])]
    raise NotImplementedError(name)


def run(n: int, ops: Sequence[tuple], noise: Dict[str, Any] = None,
        kraus_cache: Dict[str, Any] = None, gates: set = DM_GATES) -> tuple:
    """(rho, measured qubits) after the circuit (engine.py:40-102)."""
    rho = np.zeros((1 << n, 1 << n), dtype=C128)
    rho[0, 0] = 1.0
    ks = noise_kraus(noise) if noise else []
    measures = []
    for op in ops:
        name = op[0]
        if name == "measure_z":
            measures.append(int(op[1]))
        elif name in ("project_z", "reset"):
            rho = project(rho, int(op[1]), int(op[2]) if name == "project_z" else 0, n)
        elif name == "kraus":
            kops = (kraus_cache or {}).get(str(op[2]))
            if kops is not None:
                rho = apply_kraus(rho, kops, int(op[1]), n)
        elif name in gates:
            U, qs = _gate(op)
            rho = apply_unitary(rho, U, qs, n)
            for q in qs:
                if ks:
                    rho = apply_kraus(rho, ks, q, n)
    return rho, measures


def exp_z(rho: np.ndarray, q: int, n: int) -> float:  # density_matrix.py:66-73
    bits = (np.arange(1 << n) >> (n - 1 - q)) & 1
    return float(np.sum(np.real(np.diag(rho)) * (1.0 - 2.0 * bits)))


def pauli_sum_expectation(rho: np.ndarray, terms, n: int) -> float:
    """Tr(rho H) with H a Pauli term list."""
    return float(np.real(np.trace(rho @ D.pauli_sum_dense(terms, n))))
