"""Gate matrices (complex128 NumPy) with the reference's conventions.

Follows libs/quantum_library/kernels/gates.py of the reference:
  * R_P(theta) = cos(theta/2) I - i sin(theta/2) P            (gates.py:19-65, 203-252)
  * ry is the real matrix [[c, -s], [s, c]]                   (gates.py:51-65)
  * two-qubit matrices use the basis |q0 q1> with q0 the most
    significant bit; cx/cry have the control on q0            (gates.py:82-90, 265-284)
  * phase(theta) = diag(1, e^{i theta}); s = phase(pi/2),
    sd = phase(-pi/2), t = phase(pi/4), td = phase(-pi/4)     (gates.py:68-79, 187-200)

Angles are evaluated in FP64; the device receives the matrix rounded once to
complex64.  A rotation whose angle is a torch tensor that requires grad comes
back as a complex128 torch tensor built from it (cos/sin of the tensor, as the
reference builds it, gates.py:19-65), so that the kernel-level drop-ins can
differentiate through it; any other angle gives a NumPy matrix.  The
``backend`` arguments are accepted for signature compatibility.
"""
from __future__ import annotations

import math
from typing import Any, Optional, Sequence

import numpy as np

__all__ = [
    "gate_h", "gate_x", "gate_y", "gate_z", "gate_s", "gate_sd", "gate_t", "gate_td",
    "gate_phase", "gate_rx", "gate_ry", "gate_rz", "gate_cx_4x4", "gate_cz_4x4",
    "gate_cy_4x4", "gate_iswap_4x4", "gate_swap_4x4", "gate_rxx", "gate_ryy", "gate_rzz",
    "gate_cry_4x4", "gate_cx_rank4", "build_controlled_unitary", "PAULI",
]

_C = np.complex128
I2 = np.eye(2, dtype=_C)
X = np.array([[0, 1], [1, 0]], dtype=_C)
Y = np.array([[0, -1j], [1j, 0]], dtype=_C)
Z = np.array([[1, 0], [0, -1]], dtype=_C)
PAULI = {"I": I2, "X": X, "Y": Y, "Z": Z}


def _f(theta: Any) -> float:
    return float(theta.detach()) if hasattr(theta, "detach") else float(theta)


def gate_h() -> np.ndarray:
    return np.array([[1, 1], [1, -1]], dtype=_C) / math.sqrt(2.0)


def gate_x(backend: Any = None) -> np.ndarray:
    return X.copy()


def gate_y(backend: Any = None) -> np.ndarray:
    return Y.copy()


def gate_z(backend: Any = None) -> np.ndarray:
    return Z.copy()


def gate_phase(theta: Any, backend: Any = None) -> np.ndarray:
    return np.array([[1, 0], [0, np.exp(1j * _f(theta))]], dtype=_C)


def gate_s() -> np.ndarray:
    return gate_phase(math.pi / 2.0)


def gate_sd() -> np.ndarray:
    return gate_phase(-math.pi / 2.0)


def gate_t() -> np.ndarray:
    return gate_phase(math.pi / 4.0)


def gate_td() -> np.ndarray:
    return gate_phase(-math.pi / 4.0)


def _tracked(theta: Any) -> bool:
    return bool(getattr(theta, "requires_grad", False))


def _rot(P: np.ndarray, theta: Any) -> Any:
    if _tracked(theta):
        import torch

        t = theta.reshape(()).to(torch.float64)
        c = torch.cos(0.5 * t).to(torch.complex128)
        s = torch.sin(0.5 * t).to(torch.complex128)
        eye = torch.eye(P.shape[0], dtype=torch.complex128, device=t.device)
        return c * eye + (-1j * s) * torch.as_tensor(P, device=t.device)
    t = _f(theta)
    return math.cos(0.5 * t) * np.eye(P.shape[0], dtype=_C) - 1j * math.sin(0.5 * t) * P


def gate_rx(theta: Any, backend: Any = None) -> Any:
    return _rot(X, theta)


def gate_ry(theta: Any, backend: Any = None) -> Any:
    if _tracked(theta):
        import torch

        t = theta.reshape(()).to(torch.float64)
        c, s = torch.cos(0.5 * t), torch.sin(0.5 * t)
        return torch.stack([torch.stack([c, -s]), torch.stack([s, c])]).to(torch.complex128)
    t = _f(theta)
    c, s = math.cos(0.5 * t), math.sin(0.5 * t)
    return np.array([[c, -s], [s, c]], dtype=_C)


def gate_rz(theta: Any, backend: Any = None) -> Any:
    return _rot(Z, theta)


def gate_rxx(theta: Any, backend: Any = None) -> Any:
    return _rot(np.kron(X, X), theta)


def gate_ryy(theta: Any, backend: Any = None) -> Any:
    return _rot(np.kron(Y, Y), theta)


def gate_rzz(theta: Any, backend: Any = None) -> Any:
    return _rot(np.kron(Z, Z), theta)


def gate_cx_4x4() -> np.ndarray:
    return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=_C)


def gate_cx_rank4() -> np.ndarray:
    return gate_cx_4x4().reshape(2, 2, 2, 2)


def gate_cz_4x4() -> np.ndarray:
    return np.diag([1, 1, 1, -1]).astype(_C)


def gate_cy_4x4() -> np.ndarray:
    m = np.eye(4, dtype=_C)
    m[2:, 2:] = Y
    return m


def gate_iswap_4x4() -> np.ndarray:
    return np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], dtype=_C)


def gate_swap_4x4() -> np.ndarray:
    return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=_C)


def gate_cry_4x4(theta: Any, backend: Any = None) -> np.ndarray:
    m = np.eye(4, dtype=_C)
    m[2:, 2:] = gate_ry(_f(theta))
    return m


def build_controlled_unitary(U: Any, num_controls: int,
                             ctrl_state: Optional[Sequence[int]] = None) -> np.ndarray:
    """Dense multi-controlled unitary, layout [controls..., targets...].

    Same block layout as the reference's build_controlled_unitary
    (gates.py:287-318): the U block sits at the control pattern ctrl_state
    (all ones by default), identity elsewhere."""
    U = np.asarray(U, dtype=_C)
    if num_controls < 1:
        return U.copy()
    dt = U.shape[0]
    m = int(num_controls)
    cs = [1] * m if ctrl_state is None else [int(v) for v in ctrl_state]
    if len(cs) != m:
        raise ValueError("ctrl_state length must equal num_controls")
    pattern = 0
    for v in cs:
        pattern = (pattern << 1) | v
    out = np.eye((1 << m) * dt, dtype=_C)
    r = pattern * dt
    out[r:r + dt, r:r + dt] = U
    return out
