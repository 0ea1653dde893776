"""ctypes binding of oracle/sv128.c - TEST INFRASTRUCTURE ONLY.

The complex128 in-place restatement used for the full-size golden runs
(tests/golden/make_fullsize_goldens.py): the same per-gate matrices as
oracle/statevector.py (gates.py:11-318 of the reference) applied in place,
Pauli strings as (x, z) index-bit masks, and the reverse-mode gradient of
<psi|H|psi> that PyTorch autograd computes through the reference's per-gate
contractions (pytorch_backend.py:446-507 over kernels/statevector.py:28-59),
arranged as one forward and one reverse sweep:

    dE/da_k = Im <lambda_k| P_k |psi_k>,   psi_k = U_k ... U_1 |psi_0>,
    lambda_k = U_{k+1}^dagger ... U_N^dagger H psi_N,

for a gate U_k = exp(-i a_k / 2 P_k); a parameter shared by several gates
(a_k = coef_k * theta_p) collects coef_k times each term.  Pinned by
tests/test_oracle_c128.py against dense Kronecker matrices, the NumPy
oracle, the parameter-shift rule and central differences.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path
from typing import Any, List, Optional, Sequence, Tuple

import numpy as np

from . import statevector as O

HERE = Path(__file__).resolve().parent
LIB = HERE / "libsv128.so"

_lib: Optional[ctypes.CDLL] = None
_DP = ctypes.POINTER(ctypes.c_double)


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE), "-B", "-s"], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        vp, i, u64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_double
        L.sv_set_threads.argtypes = [i]
        L.sv_basis.argtypes = [vp, i, u64]
        L.sv_apply_1q.argtypes = [vp, i, i, _DP]
        L.sv_apply_2q.argtypes = [vp, i, i, i, _DP]
        L.sv_pauli_rot.argtypes = [vp, i, u64, u64, d]
        L.sv_pauli_apply_acc.argtypes = [vp, vp, i, u64, u64, d]
        L.sv_pauli_dot.argtypes = [vp, vp, i, u64, u64, _DP, _DP]
        L.sv_norm2.argtypes = [vp, i]
        L.sv_norm2.restype = d
        L.sv_copy.argtypes = [vp, vp, i]
        L.sv_zero.argtypes = [vp, i]
        _lib = L
    return _lib


def set_threads(t: int) -> None:
    lib().sv_set_threads(int(t))


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _mbuf(U: np.ndarray) -> Any:
    flat = np.asarray(U, dtype=np.complex128).reshape(-1)
    out = np.empty(2 * flat.size)
    out[0::2], out[1::2] = flat.real, flat.imag
    return out


def basis(n: int, idx: int = 0) -> np.ndarray:
    psi = np.empty(1 << n, dtype=np.complex128)
    lib().sv_basis(_p(psi), n, int(idx))
    return psi


def apply_1q(psi: np.ndarray, U: np.ndarray, q: int, n: int) -> None:
    m = _mbuf(U)
    lib().sv_apply_1q(_p(psi), n, int(q), m.ctypes.data_as(_DP))


def apply_2q(psi: np.ndarray, U4: np.ndarray, q0: int, q1: int, n: int) -> None:
    m = _mbuf(U4)
    lib().sv_apply_2q(_p(psi), n, int(q0), int(q1), m.ctypes.data_as(_DP))


def masks(paulis: Sequence[Tuple[str, int]], n: int) -> Tuple[int, int]:
    """(x, z) index-bit masks of a Pauli string [(P, qubit), ...]."""
    x = z = 0
    for P, q in paulis:
        b = 1 << (n - 1 - int(q))
        if P in ("X", "Y"):
            x ^= b
        if P in ("Z", "Y"):
            z ^= b
    return x, z


def pauli_rot(psi: np.ndarray, x: int, z: int, a: float, n: int) -> None:
    lib().sv_pauli_rot(_p(psi), n, int(x), int(z), float(a))


def pauli_dot(u: np.ndarray, v: np.ndarray, x: int, z: int, n: int) -> complex:
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().sv_pauli_dot(_p(u), _p(v), n, int(x), int(z), ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def norm2(psi: np.ndarray, n: int) -> float:
    return float(lib().sv_norm2(_p(psi), n))


def term_values(psi: np.ndarray, terms, n: int) -> np.ndarray:
    """Re <psi|P_k|psi> per term (pauli.py:74-87 semantics, coefficients not applied)."""
    return np.array([pauli_dot(psi, psi, *masks(fac, n), n).real for _, fac in terms])


def ham_apply(psi: np.ndarray, terms, n: int) -> np.ndarray:
    out = np.empty_like(psi)
    lib().sv_zero(_p(out), n)
    for c, fac in terms:
        x, z = masks(fac, n)
        lib().sv_pauli_apply_acc(_p(psi), _p(out), n, x, z, float(c))
    return out


# ---- gate lists ------------------------------------------------------------------

# (name, qubits, matrix or None, generator (x, z) or None, angle, param id, d angle/d theta)
Gate = Tuple[str, Tuple[int, ...], Optional[np.ndarray], Optional[Tuple[int, int]], float, int, float]

_ROT1 = {"rx": "X", "ry": "Y", "rz": "Z"}
_ROT2 = {"rxx": "XX", "ryy": "YY", "rzz": "ZZ"}
_FIXED1 = {"h": O.gate_h, "x": lambda: O.X, "s": O.gate_s, "sdg": O.gate_sd}
_FIXED2 = {"cx": O.gate_cx, "cz": O.gate_cz, "iswap": O.gate_iswap, "swap": O.gate_swap}


def gates_of(ops: Sequence[tuple], n: int, param_of: Optional[Sequence[Tuple[int, float]]] = None
             ) -> List[Gate]:
    """The op tuples the reference's state() dispatch applies (engine.py:914-1038
    for the kinds used by the golden workloads), as gates; rotation angles are
    floats.  param_of[k] = (param id, d angle / d theta) for the k-th rotation
    (rx/ry/rz/rxx/ryy/rzz/pauli_rot in op order), or None for no gradient."""
    out: List[Gate] = []
    r = 0
    for op in ops:
        name = op[0]
        pid, dd = -1, 0.0
        if name in _ROT1 or name in _ROT2 or name == "pauli_rot":
            if param_of is not None:
                pid, dd = param_of[r]
            r += 1
        if name in _ROT1:
            q, a = int(op[1]), float(op[2])
            out.append((name, (q,), None, masks([(_ROT1[name], q)], n), a, pid, dd))
        elif name in _ROT2:
            q0, q1, a = int(op[1]), int(op[2]), float(op[3])
            P = _ROT2[name]
            out.append((name, (q0, q1), None, masks([(P[0], q0), (P[1], q1)], n), a, pid, dd))
        elif name == "pauli_rot":
            out.append((name, (), None, masks(op[1], n), float(op[2]), pid, dd))
        elif name in _FIXED1:
            out.append((name, (int(op[1]),), _FIXED1[name](), None, 0.0, -1, 0.0))
        elif name in _FIXED2:
            out.append((name, (int(op[1]), int(op[2])), _FIXED2[name](), None, 0.0, -1, 0.0))
        elif name in ("measure_z", "barrier"):
            continue
        else:
            raise ValueError(f"op {name!r} is not handled by the complex128 C restatement")
    return out


def apply_gate(psi: np.ndarray, g: Gate, n: int, dagger: bool = False) -> None:
    name, qs, U, gen, a, _, _ = g
    if gen is not None:
        pauli_rot(psi, gen[0], gen[1], -a if dagger else a, n)
        return
    M = np.conj(U).T if dagger else U
    if len(qs) == 1:
        apply_1q(psi, M, qs[0], n)
    else:
        apply_2q(psi, M, qs[0], qs[1], n)


def run_gates(psi: np.ndarray, gates: Sequence[Gate], n: int) -> np.ndarray:
    for g in gates:
        apply_gate(psi, g, n)
    return psi


def energy_and_grad(n: int, gates: Sequence[Gate], terms, nparams: int,
                    psi0: Optional[np.ndarray] = None) -> Tuple[float, np.ndarray, float]:
    """(E, dE/dtheta[nparams], <psi|psi>) in complex128 by one forward and one
    reverse sweep (module docstring)."""
    psi = basis(n, 0) if psi0 is None else np.array(psi0, dtype=np.complex128)
    run_gates(psi, gates, n)
    nrm = norm2(psi, n)
    lam = ham_apply(psi, terms, n)
    e = pauli_dot(psi, lam, 0, 0, n).real
    grad = np.zeros(nparams)
    for g in reversed(gates):
        _, _, _, gen, _, pid, dd = g
        if pid >= 0:
            grad[pid] += dd * pauli_dot(lam, psi, gen[0], gen[1], n).imag
        apply_gate(psi, g, n, dagger=True)
        apply_gate(lam, g, n, dagger=True)
    return e, grad, nrm
