"""Density-matrix engine on the statevector kernels.

Mirrors DensityMatrixEngine (devices/simulators/density_matrix/engine.py:
33-223): ``run(circuit, shots, **kwargs)`` returning counts or <Z>
expectations, ``expval(circuit, obs)``, per-gate noise channels when
``use_noise`` is set (engine.py:185-215), Kraus ops from the circuit's
``_kraus_cache``, projection with renormalisation for project_z and reset.

The reference contracts a dense 2^n x 2^n matrix with einsum per gate and per
Kraus operator (libs/quantum_library/kernels/density_matrix.py:20-110).  Here
rho is a 2n-qubit state vec(rho) (vec[(r << n) | c] = rho[r, c]) and the
circuit becomes one fused plan on 2n qubits: a gate U on qubit q is U on
vec-qubit q (row bits) and conj(U) on vec-qubit n + q (column bits); a
channel {K_k} on qubit q is its superoperator sum_k K_k (x) conj(K_k), a 4x4
matrix on the vec-qubit pair (q, n + q).  Readouts (traces of rho P, the
diagonal for sampling) are FP64 reductions over the diagonal band
(tqsv_dm_expval, tqsv_dm_diag_amplitudes).

With ``reference_dispatch=True`` (default) only the gates the reference's
density-matrix run applies are applied (h, rz, rx, ry, cx, cz, cry, x, s,
sdg; engine.py:45-87), the others are skipped as there; with False every gate
of the statevector engine is applied.  ``expval`` traces the state that
``run`` prepares (the reference's expval rebuilds rho from h/rz/rx/cx only
and ignores noise, engine.py:153-183).
"""
from __future__ import annotations

from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from .pauli import PauliSum

DM_GATES = {"h", "rz", "rx", "ry", "cx", "cz", "cry", "x", "s", "sdg"}
_TWO_QUBIT = {"cx", "cz", "cy", "cry", "iswap", "swap", "rxx", "ryy", "rzz"}

_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_I = np.eye(2, dtype=np.complex128)


def channel_kraus(noise: Dict[str, Any]) -> List[np.ndarray]:
    """Kraus set of the per-gate noise (libs/quantum_library/noise.py:29-205,
    selected as in density_matrix/engine.py:185-215)."""
    t = str(noise.get("type", "")).lower()
    if t == "depolarizing":
        p = float(noise.get("p", 0.0))
        s = np.sqrt(p / 3.0)
        return [np.sqrt(1 - p) * _I, s * _X, s * _Y, s * _Z]
    if t == "amplitude_damping":
        g = float(noise.get("gamma", noise.get("g", 0.0)))
        return [np.array([[1, 0], [0, np.sqrt(1 - g)]], dtype=np.complex128),
                np.array([[0, np.sqrt(g)], [0, 0]], dtype=np.complex128)]
    if t == "phase_damping":
        lam = float(noise.get("lambda", noise.get("l", 0.0)))
        return [np.array([[1, 0], [0, np.sqrt(1 - lam)]], dtype=np.complex128),
                np.array([[0, 0], [0, np.sqrt(lam)]], dtype=np.complex128)]
    if t == "pauli":
        px, py, pz = (float(noise.get(k, 0.0)) for k in ("px", "py", "pz"))
        return [np.sqrt(1 - px - py - pz) * _I, np.sqrt(px) * _X, np.sqrt(py) * _Y,
                np.sqrt(pz) * _Z]
    return []


def superoperator(kops: Sequence[np.ndarray]) -> np.ndarray:
    """sum_k K (x) conj(K): the channel on (row bit, column bit), row bit the
    most significant."""
    S = np.zeros((4, 4), dtype=np.complex128)
    for K in kops:
        K = np.asarray(K, dtype=np.complex128)
        S += np.kron(K, K.conj())
    return S


def _m32(mat: np.ndarray) -> List[float]:
    flat = np.asarray(mat, dtype=np.complex128).reshape(-1)
    out = np.zeros(32)
    out[0:2 * flat.size:2] = flat.real
    out[1:2 * flat.size:2] = flat.imag
    return list(out)


def double_rows(rows: np.ndarray, n: int) -> List[np.ndarray]:
    """n-qubit tqsv_op rows -> 2n-qubit rows acting as rho -> U rho U^dag."""
    out = []
    for r in rows:
        kind = int(r["kind"])
        row = r.copy()
        col = r.copy()
        for x in (row, col):
            x["param"] = -1
            x["gx"] = 0
            x["gz"] = 0
            x["gcoef"] = 0.0
        if kind in (nat.OP_MAT1, nat.OP_MAT2):
            row["t"] = [int(t) + n if int(t) >= 0 else -1 for t in r["t"]]
            row["cmask"] = int(r["cmask"]) << n
            row["cval"] = int(r["cval"]) << n
            m = np.array(r["m"], dtype=np.float64)
            m[1::2] = -m[1::2]
            col["m"] = m
        elif kind == nat.OP_PROT:
            x, z = int(r["xmask"]), int(r["zmask"])
            row["xmask"] = x << n
            row["zmask"] = z << n
            # conj(exp(-i t/2 P)) = exp(-i t'/2 P), t' = -(-1)^{#Y} t
            ny = bin(x & z).count("1")
            col["theta"] = -float(r["theta"]) * (-1.0 if ny % 2 else 1.0)
        else:
            raise NotImplementedError(f"op kind {kind} in a density-matrix circuit")
        out += [row, col]
    return out


def _superop_row(bit: int, n: int, S: np.ndarray) -> np.ndarray:
    return np.array([(nat.OP_MAT2, (bit + n, bit), -1, 0, 0, 0, 0, 0.0, 0, 0, 0.0, _m32(S))],
                    dtype=nat.OP_DTYPE)[0]


def _wires(op: tuple) -> List[int]:
    name = op[0]
    if name in _TWO_QUBIT:
        return [int(op[1]), int(op[2])]
    if name == "pauli_rot":
        return sorted({int(q) for _, q in op[1]})
    return [int(op[1])]


class DensityMatrixEngine:
    name = "density_matrix"
    capabilities = {"supports_shots": True}

    def __init__(self, backend_name: Optional[str] = None, *, reference_dispatch: bool = True):
        from .engine import StatevectorEngine

        self._sv = StatevectorEngine(backend_name)
        self.backend = self._sv.backend
        self.reference_dispatch = reference_dispatch
        self.last_vec = None

    # -- evolution ------------------------------------------------------------
    def evolve(self, circuit: Any, noise: Optional[Dict[str, Any]] = None) -> Tuple[Any, List[int]]:
        """vec(rho) after the circuit (a 2n-qubit GPU state) and the measured qubits."""
        from .circuit import Circuit
        from .engine import get_plan, lower

        n = int(circuit.num_qubits)
        vec = nat.new_state(2 * n, 0)
        kcache = getattr(circuit, "_kraus_cache", {}) or {}
        ks = channel_kraus(noise) if noise else []
        S_noise = superoperator(ks) if ks else None
        rows: List[np.ndarray] = []
        measures: List[int] = []

        def flush():
            if rows:
                # plans cached by structure: a repeated circuit only re-emits its programs
                get_plan(2 * n, np.array(rows, dtype=nat.OP_DTYPE)).run(vec)
                rows.clear()

        for op in circuit.ops:
            if not isinstance(op, (list, tuple)) or not op:
                continue
            name = op[0]
            if name == "measure_z":
                measures.append(int(op[1]))
                continue
            if name == "barrier":
                continue
            if name in ("project_z", "reset"):
                q = int(op[1])
                keep = int(op[2]) if name == "project_z" else 0
                P = np.diag([1.0, 0.0] if keep == 0 else [0.0, 1.0]).astype(np.complex128)
                b = n - 1 - q
                rows.append(np.array([(nat.OP_MAT1, (b + n, -1), -1, 0, 0, 0, 0, 0.0, 0, 0, 0.0,
                                       _m32(P))], dtype=nat.OP_DTYPE)[0])
                rows.append(np.array([(nat.OP_MAT1, (b, -1), -1, 0, 0, 0, 0, 0.0, 0, 0, 0.0,
                                       _m32(P))], dtype=nat.OP_DTYPE)[0])
                flush()
                tr, _ = nat.dm_expval(vec, n, [0], [0])
                if abs(tr) > 0:
                    nat.scale(vec, 1.0 / tr)
                continue
            if name == "kraus":
                kops = kcache.get(str(op[2]))
                if kops is not None:
                    S = superoperator([np.asarray(k.detach().cpu().numpy() if hasattr(k, "detach")
                                                  else k, dtype=np.complex128) for k in kops])
                    rows.append(_superop_row(n - 1 - int(op[1]), n, S))
                continue
            if self.reference_dispatch and name not in DM_GATES:
                continue
            L, _, _ = lower(Circuit(n, ops=[op]), mode="run", reference_dispatch=False)
            applied = False
            for seg in L.segments:
                if seg[0] == "ops" and len(seg[1]):
                    rows.extend(double_rows(seg[1], n))
                    applied = True
            if applied and S_noise is not None:
                for q in _wires(op):
                    rows.append(_superop_row(n - 1 - q, n, S_noise))
        flush()
        self.last_vec = vec
        return vec, measures

    # -- engine contract --------------------------------------------------------
    def run(self, circuit: Any, shots: Optional[int] = None, **kwargs: Any) -> Dict[str, Any]:
        """Counts from diag(rho) for shots > 0 with measurements, else <Z_q>
        of the measured qubits (engine.py:40-151)."""
        shots = int(shots or 0)
        n = int(circuit.num_qubits)
        noise = kwargs.get("noise") if kwargs.get("use_noise") else None
        vec, measures = self.evolve(circuit, noise)
        meta = {"shots": shots, "backend": self.backend.name}
        if shots > 0 and measures:
            amp = nat.dm_diag_amplitudes(vec, n)
            return {"result": self._sv._sample(amp, n, shots, kwargs), "metadata": meta}
        x = np.zeros(len(measures), dtype=np.uint64)
        z = np.array([1 << (n - 1 - q) for q in measures], dtype=np.uint64)
        _, vals = nat.dm_expval(vec, n, x, z) if measures else (0.0, [])
        return {"expectations": {f"Z{q}": float(v) for q, v in zip(measures, vals)},
                "metadata": meta}

    def expval(self, circuit: Any, obs: Any, **kwargs: Any) -> float:
        """Tr(rho H) for a Pauli sum (term list, OpenFermion-style terms)."""
        n = int(circuit.num_qubits)
        noise = kwargs.get("noise") if kwargs.get("use_noise") else None
        vec, _ = self.evolve(circuit, noise)
        x, z, c = PauliSum.coerce(obs, n).to_masks()
        v, _ = nat.dm_expval(vec, n, x, z, c)
        return float(v)

    def density_matrix(self, circuit: Any, **kwargs: Any):
        """rho as a (2^n, 2^n) complex64 GPU tensor (a view of vec(rho))."""
        n = int(circuit.num_qubits)
        noise = kwargs.get("noise") if kwargs.get("use_noise") else None
        vec, _ = self.evolve(circuit, noise)
        return vec.view(1 << n, 1 << n)
