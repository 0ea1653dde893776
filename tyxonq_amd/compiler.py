"""The default native compile pipeline on the simulator run() path.

Restates the two passes the reference always runs before a simulator
submission (compiler/compile_engine/native/native_compiler.py:39-47):
  * rewrite/auto_measure: when the circuit has no measure_z op, append
    measure_z on every qubit and warn (compiler/stages/rewrite/auto_measure.py:23-35);
  * rewrite/gates_transform: with the default basis ["h","rx","rz","cx","cz"]
    rewrite x -> rx(pi); every other op is kept
    (compiler/stages/rewrite/gates_transform.py:24-50).
Further stages come from ``compile_plan`` (stage names as in
compiler/compile_engine/native/compile_plan.py:11-64); the measurement
stages live in tyxonq_amd.measurement.
The lowered circuit is rebuilt from (num_qubits, ops, metadata,
instructions), so, as in the reference, its unitary/kraus matrix caches and
initial state are not carried over (gates_transform.py:45-50).
"""
from __future__ import annotations

import math
import warnings
from typing import Any, Callable, Dict, Sequence, Tuple

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


def _stage(name: str) -> Callable[..., Any]:
    """Stage name -> callable(circuit, **options) (compile_plan.py:11-64).  The
    no-op stages of the reference return the circuit unchanged; stages that
    rewrite gates beyond this path (rotation decomposition, merge/prune,
    light-cone simplification) raise NotImplementedError."""
    from . import measurement as M

    identity = lambda c, **_: c
    table: Dict[str, Callable[..., Any]] = {
        "decompose": identity,
        "layout": identity,
        "scheduling": identity,
        "rewrite/auto_measure": lambda c, **_: auto_measure(c),
        "rewrite/gates_transform": lambda c, **o: gates_transform(
            c, list(DEFAULT_BASIS) if o.get("basis_gates") else None),
        "rewrite/measurement": M.rewrite_measurement,
        "scheduling/shot_scheduler": M.validate_shot_plan,
        "gradients/parameter_shift": M.parameter_shift_stage,
    }
    if name in table:
        return table[name]
    if name in ("decompose/rotations", "rewrite/merge_prune", "simplify/lightcone"):
        raise NotImplementedError(f"compile stage {name!r} is not part of tyxonq_amd")
    raise ValueError(f"Unknown stage: {name}")


class CompilePlan:
    """A sequence of stages applied in order (compile_plan.py:67-94)."""

    def __init__(self, names: Sequence[str]):
        self.names = list(names)
        self._stages = [_stage(n) for n in self.names]

    @property
    def passes(self) -> Tuple[Callable[..., Any], ...]:
        return tuple(self._stages)

    def execute_plan(self, circuit: Any, **opts: Any) -> Any:
        opts.pop("device_rule", None)
        for st in self._stages:
            circuit = st(circuit, **opts)
        return circuit


def build_plan(names: Sequence[str]) -> CompilePlan:
    return CompilePlan(names)


def compile_native(circuit: Any, options: Dict[str, Any], compile_plan: Any = None
                   ) -> Tuple[Any, Dict[str, Any]]:
    """The native pipeline: rewrite/auto_measure, rewrite/gates_transform,
    then the requested stages; a shot plan is scheduled when ``shot_plan`` or
    ``total_shots`` is given (native_compiler.py:15-70).  Returns (lowered
    circuit, metadata)."""
    from .measurement import schedule

    opts = dict(options)
    extra = compile_plan if compile_plan is not None else opts.pop("compile_plan", None)
    extra = [] if not extra else (list(extra) if isinstance(extra, (list, tuple)) else [extra])
    names = ["rewrite/auto_measure", "rewrite/gates_transform", *extra]
    lowered = build_plan(names).execute_plan(circuit, **opts)
    job_plan = None
    if "shot_plan" in opts or "total_shots" in opts:
        job_plan = schedule(lowered, opts.get("shot_plan"), total_shots=opts.get("total_shots"),
                            device_rule=opts.get("device_rule") or {})
    meta = {"options": opts, "compile_plan": names,
            "basis_gates": list(DEFAULT_BASIS) if opts.get("basis_gates") else [],
            "optimization_level": int(opts.get("optimization_level", 0)), "job_plan": job_plan}
    return lowered, meta
