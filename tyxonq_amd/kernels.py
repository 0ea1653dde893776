"""Kernel-level drop-ins for libs/quantum_library/kernels/statevector.py.

Same names, argument order and qubit convention (qubit q on index bit n-1-q)
as the reference functions (statevector.py:19-129):

    init_statevector(num_qubits, backend=None)
    apply_1q_statevector(backend, state, gate2, qubit, num_qubits)
    apply_2q_statevector(backend, state, gate4, q0, q1, num_qubits)
    apply_kqubit_unitary(state, unitary, qubit_indices, num_qubits, backend=None)
    expect_z_statevector(state, qubit, num_qubits, backend=None)

States are GPU complex64 tensors.  Like the reference (whose einsum returns a
fresh array), the apply functions return a NEW tensor and leave the input
untouched; the *_ variants work in place.  The ``backend`` arguments are
accepted for signature compatibility and ignored.
"""
from __future__ import annotations

from typing import Any, Sequence

import numpy as np

from . import _native as nat


def _as_state(state: Any, num_qubits: int):
    import torch

    if isinstance(state, torch.Tensor) and state.is_cuda and state.dtype == torch.complex64:
        t = state.reshape(-1)
    else:
        arr = state.detach() if hasattr(state, "detach") else torch.as_tensor(np.asarray(state))
        t = arr.reshape(-1).to(device=nat.device(), dtype=torch.complex64)
    if t.numel() != (1 << int(num_qubits)):
        raise ValueError(f"state has {t.numel()} amplitudes, expected 2^{num_qubits}")
    return t.contiguous()


def _mat(m: Any, dim: int) -> np.ndarray:
    if hasattr(m, "detach"):
        m = m.detach().cpu().numpy()
    return np.asarray(m, dtype=np.complex128).reshape(dim, dim)


def init_statevector(num_qubits: int, backend: Any = None):
    """|0...0> (statevector.py:19-25); for num_qubits <= 0 a one-element state."""
    import torch

    if num_qubits <= 0:
        return torch.ones(1, dtype=torch.complex64, device=nat.device())
    return nat.new_state(int(num_qubits), 0)


def apply_1q_statevector_(state, gate2: Any, qubit: int, num_qubits: int):
    psi = _as_state(state, num_qubits)
    nat.apply_1q(psi, int(num_qubits) - 1 - int(qubit), _mat(gate2, 2))
    return psi


def apply_1q_statevector(backend: Any, state: Any, gate2: Any, qubit: int, num_qubits: int):
    """statevector.py:28-42."""
    return apply_1q_statevector_(_as_state(state, num_qubits).clone(), gate2, qubit, num_qubits)


def apply_2q_statevector_(state, gate4: Any, q0: int, q1: int, num_qubits: int):
    psi = _as_state(state, num_qubits)
    if int(q0) == int(q1):
        return psi
    n = int(num_qubits)
    nat.apply_2q(psi, n - 1 - int(q0), n - 1 - int(q1), _mat(gate4, 4))
    return psi


def apply_2q_statevector(backend: Any, state: Any, gate4: Any, q0: int, q1: int, num_qubits: int):
    """statevector.py:45-59 (q0 == q1 returns the state unchanged)."""
    if int(q0) == int(q1):
        return state
    return apply_2q_statevector_(_as_state(state, num_qubits).clone(), gate4, q0, q1, num_qubits)


def apply_kqubit_unitary(state: Any, unitary: Any, qubit_indices: Sequence[int], num_qubits: int,
                         backend: Any = None):
    """statevector.py:71-129: qubit_indices[0] is the most significant index
    of the unitary; k <= 2 in place on a copy, 3 <= k <= 10 by the k-qubit
    kernel (out of place)."""
    k = len(qubit_indices)
    if k == 0:
        return state
    n = int(num_qubits)
    if k > 10:
        raise NotImplementedError("apply_kqubit_unitary supports up to 10 target qubits")
    if k >= 3:
        return nat.apply_kq(_as_state(state, n), [n - 1 - int(q) for q in qubit_indices],
                            _mat(unitary, 1 << k))
    psi = _as_state(state, n).clone()
    if k == 1:
        nat.apply_1q(psi, n - 1 - int(qubit_indices[0]), _mat(unitary, 2))
    else:
        nat.apply_2q(psi, n - 1 - int(qubit_indices[0]), n - 1 - int(qubit_indices[1]),
                     _mat(unitary, 4))
    return psi


def expect_z_statevector(state: Any, qubit: int, num_qubits: int, backend: Any = None) -> float:
    """statevector.py:62-68, accumulated in FP64."""
    psi = _as_state(state, num_qubits)
    return nat.expect_z(psi, int(num_qubits) - 1 - int(qubit))
