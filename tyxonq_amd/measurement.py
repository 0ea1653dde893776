"""Measurement feed of the simulator path: grouping, shot plans, grouped
counts energies.

Host-side mirrors of the reference pieces that sit between a Pauli sum and
the device run:

  * measurement descriptors Expectation / Probability / Sample
    (core/measurements/__init__.py:7-28);
  * product-basis grouping of Pauli sums, group_hamiltonian_pauli_terms and
    group_qubit_operator_terms (libs/hamiltonian_encoding/
    hamiltonian_grouping.py:12-150): terms with the same per-qubit basis string
    share one measurement setting, identity terms go to a constant;
  * the rewrite/measurement stage (compiler/stages/rewrite/measurement.py:
    10-160): greedy packing of measurement items whose per-wire bases agree,
    plus the Hamiltonian groups, into circuit.metadata["measurement_groups"];
  * the shot scheduler (compiler/stages/scheduling/shot_scheduler.py:47-125);
  * parameter-shift circuit pairs (compiler/gradients/parameter_shift.py:9-38).

On the GPU, ``grouped_pauli_energy`` is the counts-based energy of a Pauli sum
(the reference submits one full circuit per group, base circuit plus basis
rotations, hea_device_runtime.py:150-175): here the circuit is simulated
once, each group rotates a copy of the final state with one fused plan (H for
X, H S^dag for Y) and draws its shots with the GPU sampler, and every term of
the group is a parity sum over the distinct sampled indices.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

Bases = Tuple[str, ...]
Term = Tuple[Tuple[int, str], ...]


@dataclass(frozen=True)
class Expectation:
    """Expectation of an observable (a Pauli letter per wire) on wires."""

    obs: str
    wires: Tuple[int, ...]


@dataclass(frozen=True)
class Probability:
    wires: Tuple[int, ...]


@dataclass(frozen=True)
class Sample:
    wires: Tuple[int, ...]
    shots: int


# ---- grouping ----------------------------------------------------------------

def group_hamiltonian_pauli_terms(hamiltonian: Sequence[Tuple[float, Sequence[Tuple[str, int]]]],
                                  n_qubits: int) -> Tuple[float, Dict[Bases, List[Tuple[Term, float]]]]:
    """(identity constant, {basis string: [(term, coeff)]}) for a term list
    [(coeff, [(P, q), ...])]; basis strings are indexed by qubit, 'I' where the
    group does not measure (hamiltonian_grouping.py:128-150)."""
    const = 0.0
    groups: Dict[Bases, List[Tuple[Term, float]]] = {}
    for coeff, factors in hamiltonian:
        if not factors:
            const += float(coeff)
            continue
        term = tuple((int(q), str(P).upper()) for P, q in factors)
        key = ["I"] * int(n_qubits)
        for q, P in term:
            key[q] = P
        groups.setdefault(tuple(key), []).append((term, float(coeff)))
    return const, groups


def _real_coeff(c: Any) -> float:
    z = complex(np.asarray(c, dtype=np.complex128))
    if abs(z.imag) > 1e-10:
        raise ValueError(f"Hamiltonian term has non-negligible imaginary part: {z}")
    return float(z.real)


def group_qubit_operator_terms(qop: Any, n_qubits: int
                               ) -> Tuple[float, Dict[Bases, List[Tuple[Term, float]]]]:
    """The same grouping for an OpenFermion-style operator ({((q, P), ...): c}
    or an object with such a ``terms`` dict); complex coefficients with an
    imaginary part above 1e-10 raise ValueError (hamiltonian_grouping.py:
    97-125)."""
    terms = qop if isinstance(qop, dict) else getattr(qop, "terms", {})
    const = 0.0
    groups: Dict[Bases, List[Tuple[Term, float]]] = {}
    for key, c in terms.items():
        val = _real_coeff(c)
        if key == ():
            const += val
            continue
        term = tuple((int(q), str(P).upper()) for q, P in key)
        bases = ["I"] * int(n_qubits)
        for q, P in term:
            bases[q] = P
        groups.setdefault(tuple(bases), []).append((term, val))
    return const, groups


def _basis_map(obs: Any, wires: Tuple[int, ...]) -> Dict[int, str]:
    """Per-wire Pauli letters of a string observable ('Z', 'ZX', ...); 'I' and
    unknown forms measure nothing (measurement.py:92-117)."""
    if not isinstance(obs, str):
        return {}
    if len(obs) == len(wires):
        return {int(w): c for w, c in zip(wires, obs) if c in "XYZ"}
    if len(obs) == 1 and len(wires) == 1 and obs in ("X", "Y", "Z"):
        return {int(wires[0]): obs}
    return {}


def group_measurements(measurements: Sequence[Any]) -> List[Dict[str, Any]]:
    """Greedy first-fit packing of measurement items into groups whose
    per-wire bases agree (measurement.py:120-160)."""
    groups: List[Dict[str, Any]] = []
    for m in measurements:
        if isinstance(m, dict):
            wires = tuple(sorted(int(w) for w in m.get("wires", ())))
            obs = m.get("obs")
        else:
            wires = tuple(sorted(int(w) for w in getattr(m, "wires", ())))
            obs = getattr(m, "obs", None)
        bmap = _basis_map(obs, wires)
        for g in groups:
            if any(g["basis_map"].get(w, b) != b for w, b in bmap.items()):
                continue
            g["items"].append(m)
            g["wires"].update(wires)
            g["basis_map"].update(bmap)
            break
        else:
            groups.append({"items": [m], "wires": set(wires), "basis": "pauli",
                           "basis_map": dict(bmap)})
    for g in groups:
        g["wires"] = tuple(sorted(g["wires"]))
        g["estimated_settings"] = 1
        g["estimated_shots_per_group"] = max(1, len(g["items"])) * max(1, len(g["wires"]))
    return groups


def rewrite_measurement(circuit: Any, **opts: Any) -> Any:
    """The rewrite/measurement stage: attaches ``measurement_groups`` (and the
    identity constant of a grouped Hamiltonian) to circuit.metadata and
    returns the same circuit (measurement.py:30-89)."""
    items = list(opts.get("measurements") or [])
    if not items:
        items = [Expectation("Z", (int(op[1]),)) for op in (getattr(circuit, "ops", None) or [])
                 if isinstance(op, (list, tuple)) and len(op) >= 2
                 and str(op[0]).lower() == "measure_z"]
    groups = group_measurements(items)
    n = int(opts.get("n_qubits", getattr(circuit, "num_qubits", 0)))
    const, hgroups = 0.0, {}
    if opts.get("hamiltonian_terms") is not None:
        const, hgroups = group_hamiltonian_pauli_terms(opts["hamiltonian_terms"], n)
    elif opts.get("qubit_operator") is not None:
        const, hgroups = group_qubit_operator_terms(opts["qubit_operator"], n)
    for bases, gitems in hgroups.items():
        bmap = {q: P for q, P in enumerate(bases) if P in ("X", "Y", "Z")}
        groups.append({"items": gitems, "wires": tuple(sorted(bmap)), "basis": "pauli",
                       "basis_map": bmap, "source": "hamiltonian"})
    circuit.metadata["measurement_groups"] = groups
    if hgroups:
        circuit.metadata.setdefault("measurement_context", {})["identity_const"] = float(const)
    return circuit


def schedule(circuit: Any, shot_plan: Optional[Sequence[int]] = None, *,
             total_shots: Optional[int] = None, device_rule: Optional[Dict[str, Any]] = None
             ) -> Dict[str, Any]:
    """Shot segments: one per ``shot_plan`` entry, else ``total_shots``
    split over the measurement groups by their estimated weights (rounded,
    remainder to the last group), then split by ``max_shots_per_job`` and
    tagged with batch ids (shot_scheduler.py:47-125)."""
    segs: List[Dict[str, Any]] = []
    if shot_plan is not None:
        segs = [{"shots": int(s)} for s in shot_plan]
    else:
        groups = (getattr(circuit, "metadata", None) or {}).get("measurement_groups", [])
        if groups and total_shots is not None:
            w = [max(1, int(g.get("estimated_shots_per_group", 1))) for g in groups]
            wsum = float(sum(w))
            left = int(total_shots)
            for i, g in enumerate(groups):
                s = int(round(total_shots * (w[i] / wsum))) if i < len(groups) - 1 else left
                left -= s
                segs.append({"shots": max(0, s), "basis": g.get("basis"), "wires": g.get("wires"),
                             "basis_map": g.get("basis_map", {}), "group_index": i})
        else:
            segs = [{"shots": int(total_shots) if total_shots is not None else 0}]
    rule = dict(device_rule or {})
    cap = int(rule.get("max_shots_per_job", 0) or 0)
    if cap > 0:
        split: List[Dict[str, Any]] = []
        for seg in segs:
            left = int(seg.get("shots", 0))
            if left <= 0:
                split.append(dict(seg))
            while left > 0:
                split.append({**seg, "shots": min(cap, left)})
                left -= cap
        segs = split
    per_batch = int(rule.get("max_segments_per_batch", 0) or 0)
    if bool(rule.get("supports_batch", False)) and per_batch > 0:
        for i, seg in enumerate(segs):
            seg["batch_id"] = i // per_batch
    return {"circuit": circuit, "segments": segs}


def validate_shot_plan(circuit: Any, **opts: Any) -> Any:
    """The scheduling/shot_scheduler stage: checks ``shot_plan`` is a list of
    positive integers (shot_scheduler.py:128-141)."""
    plan = opts.get("shot_plan")
    if plan is not None and not (isinstance(plan, list)
                                 and all(isinstance(x, int) and x > 0 for x in plan)):
        raise ValueError("shot_plan must be a list of positive integers")
    return circuit


def generate_shifted_circuits(circuit: Any, op_name: str) -> Tuple[Any, Any, Dict[str, Any]]:
    """(plus, minus, {"coeff": 0.5}): copies of the circuit whose first
    ``op_name`` op with a parameter has it shifted by +-pi/2
    (compiler/gradients/parameter_shift.py:9-38)."""
    from .circuit import Circuit

    def shifted(delta: float) -> List[Any]:
        out, done = [], False
        for op in circuit.ops:
            if not done and op[0] == op_name and len(op) >= 3:
                out.append((op[0], op[1], float(op[2]) + delta) + tuple(op[3:]))
                done = True
            else:
                out.append(op)
        return out

    mk = lambda ops: Circuit(circuit.num_qubits, ops=ops, metadata=dict(circuit.metadata))
    return mk(shifted(math.pi / 2)), mk(shifted(-math.pi / 2)), {"coeff": 0.5}


def parameter_shift_stage(circuit: Any, **opts: Any) -> Any:
    """The gradients/parameter_shift stage (parameter_shift_pass.py:12-28)."""
    name = opts.get("grad_op")
    if not name:
        return circuit
    plus, minus, meta = generate_shifted_circuits(circuit, name)
    circuit.metadata.setdefault("gradients", {})[name] = {"plus": plus, "minus": minus,
                                                         "meta": meta}
    return circuit


# ---- counts energies on the GPU ---------------------------------------------

_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2.0)
_HSDG = _H @ np.diag([1.0, -1j])


# basis-rotation plans by (n, bases), kept apart from the engine's circuit
# plans (a Hamiltonian of many groups would otherwise push those out)
_ROT_PLANS: Any = None


def basis_rotation_rows(bases: Bases, n: int) -> np.ndarray:
    """Program rows that rotate the X / Y qubits of ``bases`` to Z (H, and
    H S^dag for Y, as hea_device_runtime.py:264-293 appends them)."""
    from . import _native as nat

    rows = []
    for q, P in enumerate(bases):
        if P not in ("X", "Y"):
            continue
        m = _H if P == "X" else _HSDG
        flat = np.zeros(32)
        flat[0:8:2] = m.reshape(-1).real
        flat[1:8:2] = m.reshape(-1).imag
        rows.append((nat.OP_MAT1, (n - 1 - q, -1), -1, 0, 0, 0, 0, 0.0, 0, 0, 0.0, flat))
    return np.array(rows, dtype=nat.OP_DTYPE)


def rotate_to_bases(psi: Any, bases: Bases) -> Any:
    """A copy of psi (GPU complex64) measured in ``bases``."""
    global _ROT_PLANS
    from . import _native as nat
    from .engine import _LRU

    if _ROT_PLANS is None:
        _ROT_PLANS = _LRU(512)
    n = int(psi.numel()).bit_length() - 1
    out = psi.clone()
    key = (n, tuple(bases))
    plan = _ROT_PLANS.get(key)
    if plan is None:
        rows = basis_rotation_rows(bases, n)
        plan = nat.Plan(n, rows) if len(rows) else False
        _ROT_PLANS.put(key, plan)
    if plan is not False:
        plan.run(out)
    return out


def z_products(index: np.ndarray, count: np.ndarray, terms: Sequence[Term], n: int) -> np.ndarray:
    """<prod Z_q> of each term from distinct sampled indices and their counts
    (the parity form of counts_expval.py:6-20)."""
    index = np.asarray(index, dtype=np.uint64)
    w = np.asarray(count, dtype=np.float64)
    total = max(float(w.sum()), 1.0)
    out = np.empty(len(terms))
    for k, term in enumerate(terms):
        m = 0
        for q, _P in term:
            m |= 1 << (n - 1 - int(q))
        v = index & np.uint64(m)
        par = np.zeros(index.shape, dtype=np.uint64)
        while v.any():
            par ^= v & np.uint64(1)
            v = v >> np.uint64(1)
        out[k] = float(np.dot(1.0 - 2.0 * par.astype(np.float64), w)) / total
    return out


def group_seed(seed: int, k: int) -> int:
    """Sampler seed of the k-th group: a SeedSequence child of (seed, k)."""
    return int(np.random.SeedSequence([int(seed) & (2 ** 64 - 1), int(k)])
               .generate_state(1, dtype=np.uint64)[0])


def grouped_pauli_energy(circuit: Any, hamiltonian: Any, shots: int, *, seed: Optional[int] = None,
                         engine: Any = None) -> Dict[str, Any]:
    """Counts-based <H>: identity constant + sum over product-basis groups of
    the coefficients times Z-parity averages over ``shots`` samples of the
    rotated state.  Returns {"energy", "identity_const", "groups": [{"bases",
    "items", "expvals", "index", "count"}]}.  Each group draws its shots with
    seed ``group_seed(seed, k)``."""
    from . import _native as nat
    from .engine import StatevectorEngine

    n = int(circuit.num_qubits)
    if isinstance(hamiltonian, dict) or hasattr(hamiltonian, "terms") and not isinstance(
            hamiltonian, (list, tuple)):
        const, groups = group_qubit_operator_terms(hamiltonian, n)
    else:
        const, groups = group_hamiltonian_pauli_terms(hamiltonian, n)
    if int(shots) <= 0:
        raise ValueError("shots must be positive")
    if seed is None:
        seed = int(np.random.SeedSequence().generate_state(1, dtype=np.uint64)[0])
    import torch

    eng = engine if engine is not None else StatevectorEngine(reference_dispatch=False)
    with torch.no_grad():
        psi = eng.state_tensor(circuit).detach()
    energy = float(const)
    out = []
    for k, (bases, items) in enumerate(groups.items()):
        phi = rotate_to_bases(psi, bases)
        idx, cnt = nat.sample_counts(phi, int(shots), group_seed(seed, k))
        del phi
        ev = z_products(idx, cnt, [t for t, _ in items], n)
        energy += float(np.dot(ev, [c for _, c in items]))
        out.append({"bases": bases, "items": items, "expvals": ev, "index": idx, "count": cnt})
    return {"energy": energy, "identity_const": float(const), "groups": out}
