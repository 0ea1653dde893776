"""Circuit IR: the reference's ``tq.Circuit`` surface on the GPU engine.

Follows core/ir/circuit.py of the reference: a circuit is ``num_qubits`` plus
a list of op tuples such as ("h", q), ("rx", q, theta), ("cx", c, t),
("rzz", a, b, theta), ("unitary", q[, q1], key), ("kraus", q, key[, status]),
("measure_z", q), ("project_z", q, keep), ("reset", q)
(circuit.py:1042-1632).  Builder methods append ops and return ``self``;
``device()``, ``postprocessing()`` and ``compile()`` record options;
``state()``, ``expectation()`` and ``run()`` execute on the MI355X engine.

Qubit q is bit n-1-q of the amplitude index (big-endian, engine.py:1049-1057).
"""
from __future__ import annotations

import json
import warnings
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

_GLOBAL_COMPILE_DEFAULTS: Dict[str, Any] = {}
_GLOBAL_DEVICE_DEFAULTS: Dict[str, Any] = {}
_GLOBAL_POSTPROC_DEFAULTS: Dict[str, Any] = {"method": None}


def set_global_compile_defaults(options: Dict[str, Any]) -> Dict[str, Any]:
    _GLOBAL_COMPILE_DEFAULTS.update(dict(options))
    return dict(_GLOBAL_COMPILE_DEFAULTS)


def get_global_compile_defaults() -> Dict[str, Any]:
    return dict(_GLOBAL_COMPILE_DEFAULTS)


def set_global_device_defaults(options: Dict[str, Any]) -> Dict[str, Any]:
    _GLOBAL_DEVICE_DEFAULTS.update(dict(options))
    return dict(_GLOBAL_DEVICE_DEFAULTS)


def get_global_device_defaults() -> Dict[str, Any]:
    return dict(_GLOBAL_DEVICE_DEFAULTS)


def set_global_postprocessing_defaults(options: Dict[str, Any]) -> Dict[str, Any]:
    _GLOBAL_POSTPROC_DEFAULTS.update(dict(options))
    _GLOBAL_POSTPROC_DEFAULTS.setdefault("method", None)
    return dict(_GLOBAL_POSTPROC_DEFAULTS)


def get_global_postprocessing_defaults() -> Dict[str, Any]:
    base = dict(_GLOBAL_POSTPROC_DEFAULTS)
    base.setdefault("method", None)
    return base


class Circuit:
    """Op-tuple circuit IR with chainable stages (core/ir/circuit.py:48-141)."""

    def __init__(self, num_qubits: int, ops: Optional[List[Any]] = None,
                 metadata: Optional[Dict[str, Any]] = None,
                 instructions: Optional[List[Tuple[str, Tuple[int, ...]]]] = None,
                 inputs: Optional[Any] = None,
                 compile_engine: str = "default", compile_output: str = "ir",
                 compile_target: str = "simulator::statevector",
                 compile_options: Optional[Dict[str, Any]] = None,
                 device_provider: str = "simulator", device_device: str = "statevector",
                 device_shots: int = 1024, device_options: Optional[Dict[str, Any]] = None,
                 wait_async_result: Optional[bool] = False,
                 postprocessing_method: Optional[str] = None,
                 postprocessing_options: Optional[Dict[str, Any]] = None,
                 draw_output: Optional[str] = None, source: Optional[Any] = None):
        self.num_qubits = int(num_qubits)
        self.ops: List[Any] = ops if ops is not None else []
        self.metadata: Dict[str, Any] = metadata if metadata is not None else {}
        self.instructions = instructions if instructions is not None else []
        self._initial_state = inputs
        self._compile_opts: Dict[str, Any] = get_global_compile_defaults()
        self._device_opts: Dict[str, Any] = get_global_device_defaults()
        self._post_opts: Dict[str, Any] = get_global_postprocessing_defaults()
        self._compile_engine = str(compile_engine)
        self._compile_output = str(compile_output)
        self._compile_target = str(compile_target)
        if compile_options:
            self._compile_opts.update(dict(compile_options))
        self._device_opts.setdefault("provider", str(device_provider))
        self._device_opts.setdefault("device", str(device_device))
        self._device_opts.setdefault("shots", int(device_shots))
        if device_options:
            self._device_opts.update(dict(device_options))
        self._wait_async_result = wait_async_result
        if postprocessing_options:
            self._post_opts.update(dict(postprocessing_options))
        if "method" not in self._post_opts or self._post_opts["method"] is None:
            self._post_opts["method"] = postprocessing_method
        self._compiled_source = source
        self._draw_output = str(draw_output) if draw_output is not None else None
        self.__post_init__()

    # ---- structure --------------------------------------------------------
    def __post_init__(self) -> None:
        if self.num_qubits < 0:
            raise ValueError("num_qubits must be non-negative")
        for op in self.ops:
            if not isinstance(op, (tuple, list)):
                raise TypeError("op must be a tuple or list")
            if not op:
                raise ValueError("op cannot be empty")
            if not isinstance(op[0], str):
                raise TypeError("op name must be a string")
            for arg in op[1:]:
                if isinstance(arg, int) and not isinstance(arg, bool):
                    if arg < 0 or arg >= self.num_qubits:
                        raise ValueError("qubit index out of range in op")
        for inst in self.instructions:
            if not isinstance(inst, tuple) or len(inst) != 2:
                raise TypeError("instruction must be (name, (indices,)) tuple")
            iname, idxs = inst
            if not isinstance(iname, str) or not isinstance(idxs, tuple):
                raise TypeError("instruction must be (str, tuple)")
            for q in idxs:
                if not isinstance(q, int) or q < 0 or q >= self.num_qubits:
                    raise ValueError("instruction qubit index out of range")

    def __enter__(self) -> "Circuit":
        return self

    def __exit__(self, *exc: Any) -> bool:
        return False

    def circuit(self) -> "Circuit":
        return self

    def __repr__(self) -> str:
        return f"Circuit(num_qubits={self.num_qubits}, ops={len(self.ops)})"

    def _copy_with(self, ops: List[Any]) -> "Circuit":
        c = Circuit(self.num_qubits, ops=list(ops), metadata=dict(self.metadata),
                    instructions=list(self.instructions), inputs=self._initial_state)
        c._compile_opts = dict(self._compile_opts)
        c._device_opts = dict(self._device_opts)
        c._post_opts = dict(self._post_opts)
        for attr in ("_unitary_cache", "_kraus_cache"):
            if hasattr(self, attr):
                setattr(c, attr, dict(getattr(self, attr)))
        return c

    def extended(self, extra_ops: Sequence[Any]) -> "Circuit":
        return self._copy_with(list(self.ops) + list(extra_ops))

    def with_metadata(self, **kwargs: Any) -> "Circuit":
        c = self._copy_with(self.ops)
        c.metadata.update(kwargs)
        return c

    def gate_count(self, names: Optional[Sequence[str]] = None) -> int:
        if names is None:
            return sum(1 for op in self.ops if op and op[0] not in ("measure_z", "barrier"))
        s = {str(x) for x in names}
        return sum(1 for op in self.ops if op and op[0] in s)

    def gate_summary(self) -> Dict[str, int]:
        out: Dict[str, int] = {}
        for op in self.ops:
            out[op[0]] = out.get(op[0], 0) + 1
        return out

    # ---- chainable stages ---------------------------------------------------
    def device(self, **options: Any) -> "Circuit":
        self._device_opts.update(dict(options))
        return self

    def postprocessing(self, **options: Any) -> "Circuit":
        self._post_opts.update(dict(options))
        self._post_opts.setdefault("method", None)
        return self

    def compile(self, *, compile_engine: Optional[str] = None, output: Optional[str] = None,
                target: Optional[str] = None, options: Optional[Dict[str, Any]] = None,
                **kw: Any) -> "Circuit":
        """Lower the circuit for the simulator: the default native pipeline
        rewrite/auto_measure + rewrite/gates_transform
        (compiler/compile_engine/native/native_compiler.py:39-47)."""
        from .compiler import compile_native

        opts = dict(self._compile_opts)
        if options:
            opts.update(dict(options))
        opts.update(kw)
        self._compiled_source, self._compile_meta = compile_native(self, opts)
        return self

    # ---- execution ----------------------------------------------------------
    def run(self, *, provider: Optional[str] = None, device: Optional[str] = None,
            shots: int = 1024, wait_async_result: Optional[bool] = False, **opts: Any) -> Any:
        """Compile (if needed) and execute through the device layer
        (core/ir/circuit.py:868-981).  As in the reference, the ``shots`` argument
        of run() wins over .device(shots=...) (circuit.py:873,890)."""
        from . import devices
        from .postprocessing import apply_postprocessing

        dev_opts = {**self._device_opts, **opts}
        dev_provider = provider if provider is not None else dev_opts.pop("provider", None)
        dev_device = device if device is not None else dev_opts.pop("device", None)
        dev_shots = int(shots if shots is not None else dev_opts.pop("shots", 1024))
        for k in ("provider", "device", "shots"):
            dev_opts.pop(k, None)
        if self._compiled_source is None:
            self.compile(compile_engine=self._compile_engine, output=self._compile_output,
                         target=self._compile_target, options=self._compile_opts)
        tasks = devices.run(provider=dev_provider, device=dev_device,
                            source=self._compiled_source, shots=dev_shots, **dev_opts)
        results: List[Dict[str, Any]] = []
        for t in tasks:
            rr = t.get_result(wait=bool(wait_async_result))
            has_payload = (rr.get("result") is not None) or (rr.get("expectations") is not None)
            # the run's GPU state (shots = 0): analytic post-processing reduces it in place
            psi_dev = getattr(t.handle, "device_state", None)
            if has_payload:
                rr["postprocessing"] = apply_postprocessing(rr, dict(self._post_opts),
                                                            device_state=psi_dev)
                results.append(rr)
            else:
                results.append({"result": {}, "result_meta": rr.get("result_meta", {}),
                                "postprocessing": {"method": None, "result": None}})
            if psi_dev is not None:
                t.handle.device_state = None
        # the device layer always returns a task list, so run() returns a list of result
        # dicts, as the reference does (circuit.py:979-981)
        return results

    submit_task = run

    def state(self, engine: Optional[str] = None, backend: Any = None,
              form: Optional[str] = None) -> Any:
        """Final state (core/ir/circuit.py:427-503).

        With the pytorch backend the result is the GPU complex64 tensor,
        differentiable w.r.t. tensor-valued gate angles; with the numpy backend
        (or form="numpy") a host complex128 array."""
        if engine is None:
            dev = str(self._device_opts.get("device", "statevector"))
            if "matrix_product_state" in dev or dev == "mps" or "::mps" in dev:
                engine = "mps"
            elif "density_matrix" in dev:
                engine = "density_matrix"
            else:
                engine = "statevector"
        if engine in ("mps", "matrix_product_state"):
            raise NotImplementedError("the matrix-product-state engine is not part of tyxonq_amd")
        if engine not in ("statevector", "density_matrix"):
            raise ValueError(f"Unknown engine '{engine}'. Supported: 'statevector', 'mps', "
                             "'density_matrix'")
        from .engine import StatevectorEngine

        eng = StatevectorEngine(backend_name=backend)
        out = eng.state(self)
        if form == "numpy":
            return _to_numpy(out)
        return out

    def wavefunction(self, engine: Optional[str] = None, backend: Any = None,
                     form: Optional[str] = None) -> Any:
        return self.state(engine=engine, backend=backend, form=form)

    def probability(self, backend: Any = None) -> np.ndarray:
        from .engine import StatevectorEngine

        return StatevectorEngine(backend_name=backend).probability(self)

    def amplitude(self, bitstring: str, backend: Any = None) -> complex:
        from .engine import StatevectorEngine

        return StatevectorEngine(backend_name=backend).amplitude(self, bitstring)

    def expectation(self, *pauli_ops: Any) -> Any:
        """<psi|O|psi> for a product of (matrix, qubits) factors of one or two
        qubits each (core/ir/circuit.py:1634-1796), or, given a single
        Pauli-sum argument (term list, PauliSum, {((q, P), ...): c} dict or an
        object with ``terms``), the energy of that sum in one fused pass."""
        from .engine import StatevectorEngine

        eng = StatevectorEngine()
        if len(pauli_ops) == 1 and not _is_factor(pauli_ops[0]):
            return eng.energy(self, pauli_ops[0])
        return eng.expectation_product(self, pauli_ops)

    def expval(self, *pauli_ops: Any) -> Any:
        return self.expectation(*pauli_ops)

    # ---- builder gates (core/ir/circuit.py:1042-1632) -------------------------
    def _q(self, q: Any) -> int:
        q = int(q)
        if not 0 <= q < self.num_qubits:
            raise ValueError(f"qubit index {q} out of range")
        return q

    def h(self, q: int) -> "Circuit":
        self.ops.append(("h", self._q(q)))
        return self

    def rz(self, q: int, theta: Any) -> "Circuit":
        self.ops.append(("rz", self._q(q), theta))
        return self

    def rx(self, q: int, theta: Any) -> "Circuit":
        self.ops.append(("rx", self._q(q), theta))
        return self

    def ry(self, q: int, theta: Any) -> "Circuit":
        self.ops.append(("ry", self._q(q), theta))
        return self

    def cx(self, c: int, t: int) -> "Circuit":
        self.ops.append(("cx", self._q(c), self._q(t)))
        return self

    def cz(self, c: int, t: int) -> "Circuit":
        self.ops.append(("cz", self._q(c), self._q(t)))
        return self

    def cy(self, c: int, t: int) -> "Circuit":
        self.ops.append(("cy", self._q(c), self._q(t)))
        return self

    def cry(self, c: int, t: int, theta: Any) -> "Circuit":
        self.ops.append(("cry", self._q(c), self._q(t), theta))
        return self

    def x(self, q: int) -> "Circuit":
        self.ops.append(("x", self._q(q)))
        return self

    def y(self, q: int) -> "Circuit":
        self.ops.append(("y", self._q(q)))
        return self

    def z(self, q: int) -> "Circuit":
        self.ops.append(("z", self._q(q)))
        return self

    def s(self, q: int) -> "Circuit":
        self.ops.append(("s", self._q(q)))
        return self

    def sdg(self, q: int) -> "Circuit":
        self.ops.append(("sdg", self._q(q)))
        return self

    def t(self, q: int) -> "Circuit":
        self.ops.append(("t", self._q(q)))
        return self

    def tdg(self, q: int) -> "Circuit":
        self.ops.append(("tdg", self._q(q)))
        return self

    def rxx(self, c: int, t: int, theta: Any) -> "Circuit":
        self.ops.append(("rxx", self._q(c), self._q(t), theta))
        return self

    def ryy(self, c: int, t: int, theta: Any) -> "Circuit":
        self.ops.append(("ryy", self._q(c), self._q(t), theta))
        return self

    def rzz(self, c: int, t: int, theta: Any) -> "Circuit":
        self.ops.append(("rzz", self._q(c), self._q(t), theta))
        return self

    def iswap(self, q0: int, q1: int) -> "Circuit":
        self.ops.append(("iswap", self._q(q0), self._q(q1)))
        return self

    def swap(self, q0: int, q1: int) -> "Circuit":
        self.ops.append(("swap", self._q(q0), self._q(q1)))
        return self

    def pauli_rotation(self, paulis: Sequence[Tuple[str, int]], theta: Any) -> "Circuit":
        """exp(-i theta/2 P) for a Pauli string [(P, q), ...] as one op.  The
        reference expands the same rotation into basis changes, a CX ladder
        and one rz (libs/circuits_library/utils.py:29-59)."""
        fac = tuple((str(P).upper(), self._q(q)) for P, q in paulis)
        self.ops.append(("pauli_rot", fac, theta))
        return self

    def barrier(self, *qubits: int) -> "Circuit":
        self.ops.append(("barrier",) + tuple(self._q(q) for q in qubits))
        return self

    def unitary(self, *qubits: int, matrix: Any) -> "Circuit":
        if len(qubits) == 0:
            raise ValueError("unitary requires at least one qubit index")
        if len(qubits) > 2:
            raise ValueError(f"unitary currently supports 1-2 qubits, got {len(qubits)}")
        for q in qubits:
            if not isinstance(q, (int, np.integer)) or q < 0 or q >= self.num_qubits:
                raise ValueError(f"Invalid qubit index: {q}")
        mat = _as_matrix(matrix)
        dim = 1 << len(qubits)
        if tuple(mat.shape) != (dim, dim):
            raise ValueError(f"Matrix shape {tuple(mat.shape)} incompatible with {len(qubits)} "
                             f"qubit(s), expected {(dim, dim)}")
        key = f"_unitary_{len(self.ops)}"
        if not hasattr(self, "_unitary_cache"):
            self._unitary_cache = {}
        self._unitary_cache[key] = mat
        if len(qubits) == 1:
            self.ops.append(("unitary", int(qubits[0]), key))
        else:
            self.ops.append(("unitary", int(qubits[0]), int(qubits[1]), key))
        return self

    def kraus(self, qubit: int, operators: Any, status: Optional[float] = None) -> "Circuit":
        if not isinstance(qubit, (int, np.integer)) or qubit < 0 or qubit >= self.num_qubits:
            raise ValueError(f"Invalid qubit index: {qubit}")
        key = f"_kraus_{len(self.ops)}"
        if not hasattr(self, "_kraus_cache"):
            self._kraus_cache = {}
        self._kraus_cache[key] = [_as_matrix(k) for k in operators]
        if status is not None:
            self.ops.append(("kraus", int(qubit), key, float(status)))
        else:
            self.ops.append(("kraus", int(qubit), key))
        return self

    def measure_z(self, q: int) -> "Circuit":
        self.ops.append(("measure_z", self._q(q)))
        return self

    def mid_measurement(self, q: int, keep: int = 0) -> "Circuit":
        self.ops.append(("project_z", self._q(q), int(keep)))
        return self

    def reset(self, q: int) -> "Circuit":
        warnings.warn("reset operation is typically not supported by hardware in logical circuits. "
                      "This is a simulation-only operation that projects qubit to |0> state.",
                      UserWarning, stacklevel=2)
        self.ops.append(("reset", self._q(q)))
        return self

    # upper-case aliases
    H, RZ, RX, RY, CX, CNOT, cnot, CZ, CY, CRY = h, rz, rx, ry, cx, cx, cx, cz, cy, cry
    X, Y, Z, S, Sdg, T, Tdg = x, y, z, s, sdg, t, tdg
    RXX, RYY, RZZ, ISWAP, SWAP = rxx, ryy, rzz, iswap, swap
    MEASURE_Z, RESET = measure_z, reset

    # ---- instructions / JSON ------------------------------------------------
    def add_measure(self, *qubits: int) -> "Circuit":
        c = self._copy_with(self.ops)
        c.instructions = list(self.instructions) + [("measure", (int(q),)) for q in qubits]
        return c

    def to_json_obj(self) -> Dict[str, Any]:
        ops = []
        for op in self.ops:
            ops.append([float(a) if _is_scalar(a) and not isinstance(a, (int, str)) else a
                        for a in op])
        return {"num_qubits": self.num_qubits, "ops": ops, "metadata": self.metadata,
                "instructions": [[n, list(i)] for n, i in self.instructions]}

    def to_json_str(self, indent: Optional[int] = None) -> str:
        return json.dumps(self.to_json_obj(), indent=indent)

    @classmethod
    def from_json_obj(cls, obj: Dict[str, Any]) -> "Circuit":
        inst = [(str(n), tuple(int(x) for x in idxs)) for n, idxs in obj.get("instructions", [])]
        ops = [tuple(op) for op in obj.get("ops", [])]
        return cls(num_qubits=int(obj.get("num_qubits", 0)), ops=ops,
                   metadata=dict(obj.get("metadata", {})), instructions=inst)

    @classmethod
    def from_json_str(cls, s: str) -> "Circuit":
        return cls.from_json_obj(json.loads(s))


class Hamiltonian:
    """Container for Hamiltonian terms (core/ir/circuit.py:1941-1977)."""

    def __init__(self, terms: Any):
        self.terms = terms

    def __repr__(self) -> str:
        return f"Hamiltonian(terms={self.terms!r})"


def _is_scalar(a: Any) -> bool:
    return isinstance(a, (int, float, np.floating, np.integer))


def _is_factor(obj: Any) -> bool:
    """True for a (matrix, qubits) expectation factor."""
    if not isinstance(obj, (tuple, list)) or len(obj) != 2:
        return False
    g, q = obj
    if isinstance(g, str):
        return False
    if isinstance(q, (int, np.integer)):
        return True
    return isinstance(q, (list, tuple)) and all(isinstance(v, (int, np.integer)) for v in q)


def _as_matrix(m: Any) -> np.ndarray:
    if hasattr(m, "detach"):
        m = m.detach().cpu().numpy()
    return np.asarray(m, dtype=np.complex128)


def _to_numpy(x: Any) -> np.ndarray:
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy().astype(np.complex128)
    return np.asarray(x, dtype=np.complex128)
