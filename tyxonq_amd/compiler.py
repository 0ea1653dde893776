"""The default native compile pipeline on the simulator run() path.

Restates the two passes the reference always runs before a simulator
submission (compiler/compile_engine/native/native_compiler.py:39-47):
  * rewrite/auto_measure: when the circuit has no measure_z op, append
    measure_z on every qubit and warn (compiler/stages/rewrite/auto_measure.py:23-35);
  * rewrite/gates_transform: with the default basis ["h","rx","rz","cx","cz"]
    rewrite x -> rx(pi); every other op is kept
    (compiler/stages/rewrite/gates_transform.py:24-50).
The lowered circuit is rebuilt from (num_qubits, ops, metadata,
instructions), so, as in the reference, its unitary/kraus matrix caches and
initial state are not carried over (gates_transform.py:45-50).
"""
from __future__ import annotations

import math
import warnings
from typing import Any, Dict

DEFAULT_BASIS = ["h", "rx", "rz", "cx", "cz"]


def auto_measure(circuit: Any) -> Any:
    ops = list(getattr(circuit, "ops", []) or [])
    if any(op and str(op[0]).lower() == "measure_z" for op in ops):
        return circuit
    n = int(circuit.num_qubits)
    if n <= 0:
        return circuit
    warnings.warn("No explicit measurements found; auto-added Z measurements on all qubits "
                  "during compilation.", UserWarning)
    return circuit.extended([("measure_z", q) for q in range(n)])


def gates_transform(circuit: Any, basis_gates=None) -> Any:
    from .circuit import Circuit

    basis = DEFAULT_BASIS if basis_gates is None else list(basis_gates)
    new_ops = []
    for op in getattr(circuit, "ops", []) or []:
        if not (isinstance(op, (list, tuple)) and op):
            new_ops.append(op)
            continue
        name = str(op[0]).lower()
        if name == "x" and "rx" in basis:
            new_ops.append(("rx", int(op[1]), math.pi))
        elif name == "y" and "ry" in basis:
            new_ops.append(("ry", int(op[1]), math.pi))
        else:
            new_ops.append(tuple(op))
    return Circuit(circuit.num_qubits, ops=new_ops, metadata=dict(getattr(circuit, "metadata", {})),
                   instructions=list(getattr(circuit, "instructions", [])))


def compile_native(circuit: Any, options: Dict[str, Any]) -> Any:
    basis = options.get("basis_gates") or None
    if basis:
        basis = list(DEFAULT_BASIS)
    return gates_transform(auto_measure(circuit), basis)
