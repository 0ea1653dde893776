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
untouched; the *_ variants work in place and refuse states that require grad.
The ``backend`` arguments are accepted for signature compatibility and
ignored.

Gradients flow through these functions as through the reference's einsums
("CRITICAL: ... preserve gradient chain", statevector.py:28-42): when the
state or the gate matrix requires grad, the call goes through a
torch.autograd.Function.  Its backward applies U^dagger to the incoming
gradient (one more kernel launch) and, for a gate that requires grad, forms
dL/dU[a, c] = sum_i g[i with target bits a] conj(psi[i with target bits c])
from 4^k FP64 reductions (tqsv_dot with flip and sign masks on the k target
bits, then an inverse Walsh-Hadamard transform over the sign masks).
expect_z_statevector returns a 0-dim float64 tensor on the state's device
(statevector.py:62-68 returns the backend's scalar), differentiable as well.
"""
from __future__ import annotations

from typing import Any, Sequence

import numpy as np

from . import _native as nat

MAX_GRAD_TARGETS = 4  # 4^k reductions per gate-matrix gradient


def _torch():
    import torch

    return torch


def _as_state(state: Any, num_qubits: int):
    torch = _torch()
    if isinstance(state, torch.Tensor):
        t = state.detach()
        if not (t.is_cuda and t.dtype == torch.complex64):
            t = t.to(device=nat.device(), dtype=torch.complex64)
        t = t.reshape(-1)
    else:
        t = torch.as_tensor(np.asarray(state)).reshape(-1).to(device=nat.device(),
                                                             dtype=torch.complex64)
    if t.numel() != (1 << int(num_qubits)):
        raise ValueError(f"state has {t.numel()} amplitudes, expected 2^{num_qubits}")
    return t.contiguous()


def _mat(m: Any, dim: int) -> np.ndarray:
    if hasattr(m, "detach"):
        m = m.detach().cpu().numpy()
    return np.asarray(m, dtype=np.complex128).reshape(dim, dim)


def _tracked(x: Any) -> bool:
    return bool(getattr(x, "requires_grad", False))


def _apply_new(psi, bits: Sequence[int], U: np.ndarray):
    """U on the index bits (bits[0] = most significant index of U), out of place."""
    k = len(bits)
    if k >= 3:
        return nat.apply_kq(psi, list(bits), U)
    out = psi.clone()
    if k == 1:
        nat.apply_1q(out, bits[0], U)
    else:
        nat.apply_2q(out, bits[0], bits[1], U)
    return out


def _local_mask(v: int, bits: Sequence[int]) -> int:
    """Global index mask of the local pattern v of U's index (MSB = bits[0])."""
    k = len(bits)
    m = 0
    for j, b in enumerate(bits):
        if (v >> (k - 1 - j)) & 1:
            m |= 1 << int(b)
    return m


def _matrix_grad(psi, g, bits: Sequence[int]) -> np.ndarray:
    """G[a, c] = sum over i with target bits a of g[i] conj(psi[i^a^c]).

    tqsv_dot(psi, g, x, z) = sum_i (-1)^{|i & z|} conj(psi[i]) g[i ^ x]
    = sum_c (-1)^{c.z} G[c ^ d, c] for the local flip pattern d of x; the
    inverse Walsh-Hadamard transform over z gives G[c ^ d, c] for every c."""
    k = len(bits)
    dim = 1 << k
    G = np.zeros((dim, dim), dtype=np.complex128)
    for d in range(dim):
        xm = _local_mask(d, bits)
        D = np.array([nat.dot(psi, g, xm, _local_mask(z, bits)) for z in range(dim)])
        for c in range(dim):
            sgn = np.array([(-1) ** bin(c & z).count("1") for z in range(dim)], dtype=float)
            G[c ^ d, c] = np.dot(sgn, D) / dim
    return G


def _apply_function():
    torch = _torch()

    class _ApplyUnitary(torch.autograd.Function):
        """out = U psi on the target bits; backward: U^dagger g and g psi^dagger."""

        @staticmethod
        def forward(ctx, state, gate, n, bits):
            psi = _as_state(state, n)
            U = _mat(gate, 1 << len(bits))
            out = _apply_new(psi, bits, U)
            ctx.U = U
            ctx.bits = bits
            ctx.n = n
            ctx.state_meta = (state.dtype, state.device, state.shape) \
                if isinstance(state, torch.Tensor) else None
            ctx.gate_meta = (gate.dtype, gate.device, gate.shape) \
                if isinstance(gate, torch.Tensor) else None
            if ctx.needs_input_grad[1]:
                if len(bits) > MAX_GRAD_TARGETS:
                    raise NotImplementedError(
                        f"gradients w.r.t. a unitary on more than {MAX_GRAD_TARGETS} qubits")
                ctx.save_for_backward(psi)
            return out

        @staticmethod
        def backward(ctx, gout):
            g = gout.detach().to(device=nat.device(), dtype=torch.complex64).reshape(-1) \
                .contiguous()
            gs = gU = None
            if ctx.needs_input_grad[0]:
                gs = _apply_new(g, ctx.bits, ctx.U.conj().T)
                dt, dev, shape = ctx.state_meta
                gs = gs.to(device=dev, dtype=dt if dt.is_complex else torch.complex128)
                gs = gs.reshape(shape)
            if ctx.needs_input_grad[1]:
                (psi,) = ctx.saved_tensors
                G = _matrix_grad(psi, g, ctx.bits)
                dt, dev, shape = ctx.gate_meta
                gU = torch.as_tensor(G, dtype=torch.complex128).to(device=dev, dtype=dt)
                gU = gU.reshape(shape)
            return gs, gU, None, None

    return _ApplyUnitary


def _expect_z_function():
    torch = _torch()

    class _ExpectZ(torch.autograd.Function):
        """<Z_bit> in FP64; d<Z>/dpsi (torch's convention) = 2 Z psi."""

        @staticmethod
        def forward(ctx, state, n, bit):
            psi = _as_state(state, n)
            val = nat.expect_z(psi, bit)
            ctx.bit = bit
            ctx.state_meta = (state.dtype, state.device, state.shape)
            ctx.save_for_backward(psi)
            return torch.tensor(val, dtype=torch.float64, device=state.device)

        @staticmethod
        def backward(ctx, ge):
            (psi,) = ctx.saved_tensors
            zpsi = psi.clone()
            nat.apply_1q(zpsi, ctx.bit, np.diag([1.0, -1.0]).astype(np.complex128))
            nat.scale(zpsi, 2.0 * float(ge))
            dt, dev, shape = ctx.state_meta
            return zpsi.to(device=dev, dtype=dt if dt.is_complex else torch.complex128) \
                .reshape(shape), None, None

    return _ExpectZ


_FUNCS: dict = {}


def _fn(name: str):
    f = _FUNCS.get(name)
    if f is None:
        f = _FUNCS[name] = {"apply": _apply_function, "expz": _expect_z_function}[name]()
    return f


def _apply(state: Any, gate: Any, bits: Sequence[int], n: int):
    if _tracked(state) or _tracked(gate):
        return _fn("apply").apply(state, gate, n, tuple(int(b) for b in bits))
    return _apply_new(_as_state(state, n), list(bits), _mat(gate, 1 << len(bits)))


def _inplace_state(state: Any, num_qubits: int):
    if _tracked(state):
        raise ValueError("in-place kernels cannot take a state that requires grad; "
                         "use the out-of-place function")
    return _as_state(state, num_qubits)


def init_statevector(num_qubits: int, backend: Any = None):
    """|0...0> (statevector.py:19-25); for num_qubits <= 0 a one-element state."""
    torch = _torch()
    if num_qubits <= 0:
        return torch.ones(1, dtype=torch.complex64, device=nat.device())
    return nat.new_state(int(num_qubits), 0)


def apply_1q_statevector_(state, gate2: Any, qubit: int, num_qubits: int):
    psi = _inplace_state(state, num_qubits)
    nat.apply_1q(psi, int(num_qubits) - 1 - int(qubit), _mat(gate2, 2))
    return psi


def apply_1q_statevector(backend: Any, state: Any, gate2: Any, qubit: int, num_qubits: int):
    """statevector.py:28-42."""
    n = int(num_qubits)
    return _apply(state, gate2, [n - 1 - int(qubit)], n)


def apply_2q_statevector_(state, gate4: Any, q0: int, q1: int, num_qubits: int):
    psi = _inplace_state(state, num_qubits)
    if int(q0) == int(q1):
        return psi
    n = int(num_qubits)
    nat.apply_2q(psi, n - 1 - int(q0), n - 1 - int(q1), _mat(gate4, 4))
    return psi


def apply_2q_statevector(backend: Any, state: Any, gate4: Any, q0: int, q1: int, num_qubits: int):
    """statevector.py:45-59 (q0 == q1 returns the state unchanged)."""
    if int(q0) == int(q1):
        return state
    n = int(num_qubits)
    return _apply(state, gate4, [n - 1 - int(q0), n - 1 - int(q1)], n)


def apply_kqubit_unitary(state: Any, unitary: Any, qubit_indices: Sequence[int], num_qubits: int,
                         backend: Any = None):
    """statevector.py:71-129: qubit_indices[0] is the most significant index
    of the unitary; k <= 2 by the 1q/2q kernels on a copy, 3 <= k <= 10 by the
    k-qubit kernel (out of place)."""
    k = len(qubit_indices)
    if k == 0:
        return state
    n = int(num_qubits)
    if k > 10:
        raise NotImplementedError("apply_kqubit_unitary supports up to 10 target qubits")
    if len(set(int(q) for q in qubit_indices)) != k:
        raise ValueError("qubit_indices must be distinct")
    return _apply(state, unitary, [n - 1 - int(q) for q in qubit_indices], n)


def expect_z_statevector(state: Any, qubit: int, num_qubits: int, backend: Any = None):
    """statevector.py:62-68, accumulated in FP64; a 0-dim float64 tensor on
    the state's device (differentiable when the state requires grad)."""
    torch = _torch()
    n = int(num_qubits)
    bit = n - 1 - int(qubit)
    if _tracked(state):
        return _fn("expz").apply(state, n, bit)
    psi = _as_state(state, n)
    dev = state.device if isinstance(state, torch.Tensor) else psi.device
    return torch.tensor(nat.expect_z(psi, bit), dtype=torch.float64, device=dev)


def parameter_shift_gradient(energy_fn: Any, params: Sequence[float]) -> np.ndarray:
    """g_i = (E(theta + pi/2 e_i) - E(theta - pi/2 e_i)) / 2 for a 1-D parameter
    vector (kernels/common.py:11-23): exact for gates generated by Pauli
    operators with eigenvalues +-1/2.  energy_fn is called 2 * len(params)
    times; the adjoint sweep (StatevectorEngine.energy_and_grad) gives the same
    gradient in one forward and one reverse pass."""
    base = np.asarray(params, dtype=np.float64)
    if base.ndim != 1:
        raise ValueError("params must be 1-D")
    s = 0.5 * np.pi
    g = np.zeros_like(base)
    for i in range(base.size):
        p_plus = base.copy()
        p_plus[i] += s
        p_minus = base.copy()
        p_minus[i] -= s
        g[i] = 0.5 * (float(energy_fn(p_plus)) - float(energy_fn(p_minus)))
    return g
