"""NumPy complex128 restatement of the reference's statevector path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Every function names the
reference file:line (under src/tyxonq/) whose algorithm it restates.  The
reference builds its einsum spec from 26 letters and fails for n >= 23
(2-qubit gates); the restatement contracts with tensordot, which is the
same contraction without the letter limit.
"""
from __future__ import annotations

import math
from typing import Any, Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

C128 = np.complex128

# ---- gate matrices: libs/quantum_library/kernels/gates.py -------------------
X = np.array([[0, 1], [1, 0]], dtype=C128)           # gates.py:163-170
Y = np.array([[0, -1j], [1j, 0]], dtype=C128)        # gates.py:173-177
Z = np.array([[1, 0], [0, -1]], dtype=C128)          # gates.py:180-184
I2 = np.eye(2, dtype=C128)
PAULI = {"I": I2, "X": X, "Y": Y, "Z": Z}


def gate_h() -> np.ndarray:  # gates.py:11-16
    return np.array([[1, 1], [1, -1]], dtype=C128) / math.sqrt(2.0)


def gate_rz(t: float) -> np.ndarray:  # gates.py:19-35: cos(t/2) I - i sin(t/2) Z
    return math.cos(t / 2) * I2 - 1j * math.sin(t / 2) * Z


def gate_rx(t: float) -> np.ndarray:  # gates.py:38-48
    return math.cos(t / 2) * I2 - 1j * math.sin(t / 2) * X


def gate_ry(t: float) -> np.ndarray:  # gates.py:51-65: [[c, -s], [s, c]]
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=C128)


def gate_phase(t: float) -> np.ndarray:  # gates.py:68-79
    return np.array([[1, 0], [0, np.exp(1j * t)]], dtype=C128)


def gate_s() -> np.ndarray:  # gates.py:187-188
    return gate_phase(math.pi / 2)


def gate_sd() -> np.ndarray:  # gates.py:191-192
    return gate_phase(-math.pi / 2)


def gate_cx() -> np.ndarray:  # gates.py:82-90 (control = first qubit, basis |q0 q1>)
    return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=C128)


def gate_cz() -> np.ndarray:  # gates.py:98-107
    return np.diag([1, 1, 1, -1]).astype(C128)


def gate_iswap() -> np.ndarray:  # gates.py:110-133
    return np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], dtype=C128)


def gate_swap() -> np.ndarray:  # gates.py:136-160
    return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=C128)


def gate_rxx(t: float) -> np.ndarray:  # gates.py:203-215
    return math.cos(t / 2) * np.eye(4, dtype=C128) - 1j * math.sin(t / 2) * np.kron(X, X)


def gate_ryy(t: float) -> np.ndarray:  # gates.py:218-235
    return math.cos(t / 2) * np.eye(4, dtype=C128) - 1j * math.sin(t / 2) * np.kron(Y, Y)


def gate_rzz(t: float) -> np.ndarray:  # gates.py:238-252
    return math.cos(t / 2) * np.eye(4, dtype=C128) - 1j * math.sin(t / 2) * np.kron(Z, Z)


def gate_cry(t: float) -> np.ndarray:  # gates.py:265-284 (control = first qubit)
    m = np.eye(4, dtype=C128)
    m[2:, 2:] = gate_ry(t)
    return m


# ---- kernels: libs/quantum_library/kernels/statevector.py -------------------

def init_statevector(n: int) -> np.ndarray:
    """statevector.py:19-25: |0...0>, complex128."""
    psi = np.zeros(1 << n, dtype=C128)
    psi[0] = 1.0
    return psi


def apply_1q(psi: np.ndarray, U: np.ndarray, q: int, n: int) -> np.ndarray:
    """statevector.py:28-42: contract U[a, b] with axis q of the (2,)*n tensor,
    the new axis a replacing axis q (qubit q is axis q, big-endian)."""
    t = np.asarray(psi, dtype=C128).reshape((2,) * n)
    out = np.tensordot(np.asarray(U, dtype=C128), t, axes=([1], [q]))
    return np.moveaxis(out, 0, q).reshape(-1)


def apply_2q(psi: np.ndarray, U4: np.ndarray, q0: int, q1: int, n: int) -> np.ndarray:
    """statevector.py:45-59: U reshaped (2,2,2,2) as [out q0, out q1, in q0, in q1];
    q0 == q1 returns the state unchanged (statevector.py:46-47)."""
    if q0 == q1:
        return psi
    t = np.asarray(psi, dtype=C128).reshape((2,) * n)
    U = np.asarray(U4, dtype=C128).reshape(2, 2, 2, 2)
    out = np.tensordot(U, t, axes=([2, 3], [q0, q1]))
    return np.moveaxis(out, [0, 1], [q0, q1]).reshape(-1)


def apply_kqubit(psi: np.ndarray, U: np.ndarray, qubits: Sequence[int], n: int) -> np.ndarray:
    """statevector.py:71-129: move the target axes last (in the given order),
    psi_matrix @ U^T, inverse permutation."""
    k = len(qubits)
    if k == 0:
        return psi
    keep = [a for a in range(n) if a not in qubits]
    perm = keep + list(qubits)
    inv = np.argsort(perm)
    t = np.asarray(psi, dtype=C128).reshape((2,) * n).transpose(perm).reshape(1 << (n - k), 1 << k)
    out = t @ np.asarray(U, dtype=C128).reshape(1 << k, 1 << k).T
    return out.reshape((2,) * n).transpose(inv).reshape(-1)


def expect_z(psi: np.ndarray, q: int, n: int) -> float:
    """statevector.py:62-68: sum |psi|^2 over qubit q = 0 minus over q = 1."""
    s = np.moveaxis(np.asarray(psi).reshape((2,) * n), q, 0).reshape(2, -1)
    p = np.sum(np.abs(s) ** 2, axis=1)
    return float(p[0] - p[1])


def project_z(psi: np.ndarray, q: int, keep: int, n: int) -> np.ndarray:
    """engine.py:1075-1087: zero the other half, renormalise if the norm is > 0."""
    t = np.moveaxis(np.array(psi, dtype=C128).reshape((2,) * n), q, 0)
    t[1 - keep, ...] = 0
    out = np.moveaxis(t, 0, q).reshape(-1)
    nrm = np.linalg.norm(out)
    return out / nrm if nrm > 0 else out


def apply_kraus(psi: np.ndarray, kops: Sequence[np.ndarray], q: int, n: int,
                status: Optional[float]) -> np.ndarray:
    """statevector.py:132-218: pick K_i by the Born weights ||K_i psi||^2 and the
    cumulative threshold `status`, apply it, renormalise."""
    outs = [apply_1q(psi, K, q, n) for K in kops]
    probs = np.array([np.real(np.vdot(o, o)) for o in outs])
    probs = probs / probs.sum()
    if status is None:
        status = np.random.random()
    sel = 0
    for i, c in enumerate(np.cumsum(probs)):
        if status <= float(c):
            sel = i
            break
    o = outs[sel]
    return o / np.sqrt(np.real(np.vdot(o, o)))


# ---- engine dispatch: devices/simulators/statevector/engine.py -----------------

_STATE_SKIP = {"y", "z", "t", "tdg", "cy", "cry"}   # not dispatched by state(), engine.py:914-1038
_RUN_SKIP = {"y", "z", "t", "tdg", "cy"}           # not dispatched by run(), engine.py:52-374


def _f(v: Any) -> float:
    return float(v.detach()) if hasattr(v, "detach") else float(v)


def simulate(n: int, ops: Sequence[tuple], *, mode: str = "state", init: Optional[np.ndarray] = None,
             unitary_cache: Optional[Dict[str, Any]] = None,
             kraus_cache: Optional[Dict[str, Any]] = None,
             reference_dispatch: bool = True) -> Tuple[np.ndarray, List[int]]:
    """Op dispatch of StatevectorEngine.state (engine.py:897-1039) / run
    (engine.py:43-374): returns (final state, measured qubits)."""
    psi = init_statevector(n) if init is None else np.asarray(init, dtype=C128).reshape(-1).copy()
    skip = (_STATE_SKIP if mode == "state" else _RUN_SKIP) if reference_dispatch else set()
    ucache = unitary_cache or {}
    kcache = kraus_cache or {}
    meas: List[int] = []
    for op in ops:
        name = op[0]
        if name in skip:
            continue
        if name == "h":
            psi = apply_1q(psi, gate_h(), op[1], n)
        elif name == "rz":
            psi = apply_1q(psi, gate_rz(_f(op[2])), op[1], n)
        elif name == "rx":
            psi = apply_1q(psi, gate_rx(_f(op[2])), op[1], n)
        elif name == "ry":
            psi = apply_1q(psi, gate_ry(_f(op[2])), op[1], n)
        elif name == "cx":
            psi = apply_2q(psi, gate_cx(), op[1], op[2], n)
        elif name == "cry":
            psi = apply_2q(psi, gate_cry(_f(op[3])), op[1], op[2], n)
        elif name == "cz":
            psi = apply_2q(psi, gate_cz(), op[1], op[2], n)
        elif name == "iswap":
            psi = apply_2q(psi, gate_iswap(), op[1], op[2], n)
        elif name == "swap":
            psi = apply_2q(psi, gate_swap(), op[1], op[2], n)
        elif name == "rxx":
            psi = apply_2q(psi, gate_rxx(_f(op[3])), op[1], op[2], n)
        elif name == "ryy":
            psi = apply_2q(psi, gate_ryy(_f(op[3])), op[1], op[2], n)
        elif name == "rzz":
            psi = apply_2q(psi, gate_rzz(_f(op[3])), op[1], op[2], n)
        elif name == "x":
            psi = apply_1q(psi, X, op[1], n)
        elif name == "y":
            psi = apply_1q(psi, Y, op[1], n)
        elif name == "z":
            psi = apply_1q(psi, Z, op[1], n)
        elif name == "s":
            psi = apply_1q(psi, gate_s(), op[1], n)
        elif name == "sdg":
            psi = apply_1q(psi, gate_sd(), op[1], n)
        elif name == "t":
            psi = apply_1q(psi, gate_phase(math.pi / 4), op[1], n)
        elif name == "tdg":
            psi = apply_1q(psi, gate_phase(-math.pi / 4), op[1], n)
        elif name == "cy":
            m = np.eye(4, dtype=C128)
            m[2:, 2:] = Y
            psi = apply_2q(psi, m, op[1], op[2], n)
        elif name == "pauli_rot":
            psi = pauli_rotation(psi, op[1], _f(op[2]), n)
        elif name == "measure_z":
            meas.append(int(op[1]))
        elif name == "project_z":
            psi = project_z(psi, int(op[1]), int(op[2]), n)
        elif name == "reset":
            psi = project_z(psi, int(op[1]), 0, n)
        elif name == "unitary":
            mat = ucache.get(str(op[-1]))
            if mat is None:
                continue
            qs = [int(a) for a in op[1:-1]]
            psi = apply_kqubit(psi, np.asarray(mat, dtype=C128), qs, n)
        elif name == "kraus":
            kops = kcache.get(str(op[2]))
            if kops is None:
                continue
            status = float(op[3]) if len(op) > 3 else None
            psi = apply_kraus(psi, [np.asarray(k, dtype=C128) for k in kops], int(op[1]), n, status)
    return psi, meas


# ---- Pauli strings and expectations -------------------------------------------

def pauli_apply(psi: np.ndarray, paulis: Sequence[Tuple[str, int]], n: int) -> np.ndarray:
    for P, q in paulis:
        psi = apply_1q(psi, PAULI[P], int(q), n)
    return psi


def expectation_product(psi: np.ndarray, factors: Sequence[Tuple[np.ndarray, Sequence[int]]],
                        n: int) -> float:
    """core/ir/circuit.py:1751-1796: apply each (matrix, qubits) factor to a
    copy, then Re tensordot(conj(psi'), psi)."""
    phi = np.array(psi, dtype=C128)
    for m, qs in factors:
        if len(qs) == 1:
            phi = apply_1q(phi, m, qs[0], n)
        elif len(qs) == 2:
            phi = apply_2q(phi, m, qs[0], qs[1], n)
        else:
            raise NotImplementedError
    return float(np.real(np.tensordot(np.conj(phi), psi, axes=1)))


def pauli_sum_expectation(psi: np.ndarray, terms: Sequence[Tuple[float, Sequence[Tuple[str, int]]]],
                          n: int, per_term: bool = False):
    """sum_k c_k <psi|P_k|psi> by applying each Pauli string to psi."""
    vals = []
    for c, fac in terms:
        phi = pauli_apply(psi, fac, n)
        vals.append(float(np.real(np.vdot(psi, phi))))
    e = float(sum(float(c) * v for (c, _), v in zip(terms, vals)))
    return (e, vals) if per_term else e


def pauli_rotation(psi: np.ndarray, paulis: Sequence[Tuple[str, int]], theta: float,
                   n: int) -> np.ndarray:
    """exp(-i theta/2 P) psi = cos(theta/2) psi - i sin(theta/2) P psi, the
    rotation evolve_pauli_ops builds from gates (libs/circuits_library/utils.py:29-59)."""
    return math.cos(theta / 2) * psi - 1j * math.sin(theta / 2) * pauli_apply(psi, paulis, n)


# ---- gradients ---------------------------------------------------------------

def parameter_shift_gradient(energy_fn: Callable[[np.ndarray], float],
                             params: Sequence[float]) -> np.ndarray:
    """libs/quantum_library/kernels/common.py:11-23: g_i = (E(+pi/2) - E(-pi/2))/2."""
    base = np.asarray(params, dtype=np.float64)
    g = np.zeros_like(base)
    for i in range(len(base)):
        p = base.copy()
        p[i] += 0.5 * np.pi
        m = base.copy()
        m[i] -= 0.5 * np.pi
        g[i] = 0.5 * (float(energy_fn(p)) - float(energy_fn(m)))
    return g


def central_fd_gradient(energy_fn: Callable[[np.ndarray], float], params: Sequence[float],
                        h: float = 1e-3, indices: Optional[Sequence[int]] = None) -> np.ndarray:
    """Central finite differences in FP64 (the FD fallback of
    numerics/backends/pytorch_backend.py:520-560 uses the same stencil)."""
    base = np.asarray(params, dtype=np.float64)
    idx = range(len(base)) if indices is None else indices
    g = np.zeros(len(list(idx)) if indices is not None else len(base))
    for j, i in enumerate(range(len(base)) if indices is None else indices):
        p = base.copy()
        p[i] += h
        m = base.copy()
        m[i] -= h
        g[j] = (float(energy_fn(p)) - float(energy_fn(m))) / (2 * h)
    return g


def torch_autograd_energy(n: int, build_ops: Callable[[Any], Sequence[tuple]],
                          terms: Sequence[Tuple[float, Sequence[Tuple[str, int]]]],
                          params: Sequence[float]) -> Tuple[float, np.ndarray]:
    """Energy and gradient by PyTorch-CPU autograd through the same per-gate
    contractions, complex128: the gradient path of the reference's pytorch
    backend (pytorch_backend.py:446-507 through kernels/statevector.py:28-59)."""
    import torch

    th = torch.tensor(np.asarray(params, dtype=np.float64), requires_grad=True)
    ops = build_ops(th)

    def t1(psi, U, q):
        t = psi.reshape((2,) * n)
        out = torch.tensordot(U, t, dims=([1], [q]))
        return torch.movedim(out, 0, q).reshape(-1)

    def t2(psi, U4, q0, q1):
        t = psi.reshape((2,) * n)
        out = torch.tensordot(U4.reshape(2, 2, 2, 2), t, dims=([2, 3], [q0, q1]))
        return torch.movedim(out, (0, 1), (q0, q1)).reshape(-1)

    def rot(P, a):
        c = torch.cos(a / 2).to(torch.complex128)
        s = torch.sin(a / 2).to(torch.complex128)
        return c * torch.eye(P.shape[0], dtype=torch.complex128) - 1j * s * P

    tX, tY, tZ = (torch.tensor(m) for m in (X, Y, Z))
    psi = torch.zeros(1 << n, dtype=torch.complex128)
    psi[0] = 1
    for op in ops:
        name = op[0]
        if name == "rx":
            psi = t1(psi, rot(tX, op[2]), op[1])
        elif name == "ry":
            c = torch.cos(op[2] / 2)
            s = torch.sin(op[2] / 2)
            U = torch.stack([torch.stack([c, -s]), torch.stack([s, c])]).to(torch.complex128)
            psi = t1(psi, U, op[1])
        elif name == "rz":
            psi = t1(psi, rot(tZ, op[2]), op[1])
        elif name == "h":
            psi = t1(psi, torch.tensor(gate_h()), op[1])
        elif name == "cx":
            psi = t2(psi, torch.tensor(gate_cx()), op[1], op[2])
        elif name == "cz":
            psi = t2(psi, torch.tensor(gate_cz()), op[1], op[2])
        elif name == "pauli_rot":
            phi = psi
            for P, q in op[1]:
                phi = t1(phi, {"X": tX, "Y": tY, "Z": tZ}[P], q)
            a = op[2]
            psi = torch.cos(a / 2) * psi - 1j * torch.sin(a / 2) * phi
        else:
            raise ValueError(f"op {name} not supported by the autograd restatement")
    e = torch.zeros((), dtype=torch.float64)
    for c, fac in terms:
        phi = psi
        for P, q in fac:
            phi = t1(phi, {"X": tX, "Y": tY, "Z": tZ}[P], q)
        e = e + float(c) * torch.real(torch.vdot(psi, phi))
    e.backward()
    return float(e.detach()), th.grad.detach().numpy().copy()


# ---- run() expectations with the attenuation noise model ----------------------

_ATTEN_1W = {"h", "rz", "rx", "ry", "x", "s", "sdg"}                       # engine.py:56-71,100-111
_ATTEN_2W = {"cx", "cry", "cz", "iswap", "swap", "rxx", "ryy", "rzz"}        # engine.py:72-99


def attenuated_wires(op: Sequence[Any], unitary_cache: Optional[Dict[str, Any]] = None
                     ) -> List[int]:
    """The qubits whose <Z> run() attenuates after op: the op's named wires
    only, as in the per-op _attenuate(noise, z_atten, [q]) / [c, t] calls of
    engine.py:56-141 (never its angle); a unitary op counts only when its
    matrix is in the cache and it acts on one or two qubits (engine.py:124-141)."""
    name = op[0]
    if name in _ATTEN_1W:
        return [int(op[1])]
    if name in _ATTEN_2W:
        return [int(op[1]), int(op[2])]
    if name == "unitary" and len(op) in (3, 4) and (unitary_cache or {}).get(str(op[-1])) is not None:
        return [int(a) for a in op[1:-1]]
    return []


def run_expectations(n: int, ops: Sequence[tuple], noise: Optional[Dict[str, Any]] = None,
                     unitary_cache: Optional[Dict[str, Any]] = None) -> Dict[str, float]:
    """StatevectorEngine.run with shots=0 (engine.py:43-473): <Z_q> for every
    measure_z qubit, each scaled by prod over the gates touching q of
    max(0, 1 - 4p/3) under depolarizing noise (engine.py:486-494)."""
    psi, meas = simulate(n, ops, mode="run", unitary_cache=unitary_cache)
    fac = 1.0
    if noise is not None and str(noise.get("type", "")).lower() == "depolarizing":
        fac = max(0.0, 1.0 - 4.0 * float(noise.get("p", 0.0)) / 3.0)
    atten = [1.0] * n
    if noise is not None:
        for op in ops:
            for q in attenuated_wires(op, unitary_cache):
                atten[q] *= fac
    return {f"Z{q}": expect_z(psi, q, n) * atten[q] for q in meas}


def build_controlled_unitary(U: np.ndarray, num_controls: int,
                             ctrl_state: Optional[Sequence[int]] = None) -> np.ndarray:
    """gates.py:287-318: block layout [controls..., targets...]; the block whose
    control pattern matches ctrl_state (bit i of the block index equals
    ctrl_state[m-1-i]) holds U, every other block the identity."""
    U = np.asarray(U, dtype=C128)
    m = int(num_controls)
    if m < 1:
        return U
    dt = U.shape[0]
    cs = [1] * m if ctrl_state is None else list(ctrl_state)
    out = np.zeros(((1 << m) * dt, (1 << m) * dt), dtype=C128)
    for mask in range(1 << m):
        row = mask * dt
        if all(((mask >> i) & 1) == cs[m - 1 - i] for i in range(m)):
            out[row:row + dt, row:row + dt] = U
        else:
            out[row:row + dt, row:row + dt] = np.eye(dt, dtype=C128)
    return out
