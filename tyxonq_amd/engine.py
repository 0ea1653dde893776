"""StatevectorEngine: the reference's simulator engine on the MI355X.

Replaces devices/simulators/statevector/engine.py of the reference.  The
public surface is the same duck-typed engine contract that the simulator
driver and ``Circuit`` call (engine.py:35-1072):

  * ``StatevectorEngine(backend_name=None)``
  * ``run(circuit, shots=None, **kwargs)`` -> {"result": counts, "metadata"}
    for shots > 0 with measurements, else {"expectations": {"Z{q}": v}, ...}
  * ``state(circuit)``, ``expval(circuit, obs)``, ``probability(circuit)``,
    ``amplitude(circuit, bitstring)``, ``perfect_sampling(circuit, rng=None)``

Execution: the circuit's op tuples are lowered to a flat list of ``tqsv_op``
records (bit = n-1-qubit), planned into fused tile passes by libtqsv, and run
in place on a complex64 GPU state.  Measurement-like ops (project_z, reset,
kraus) split the list into segments.

Dispatch follows the reference op by op, including what it skips: in
``state()`` the ops y, z, t, tdg, cy and cry are ignored (engine.py:914-1038)
and in ``run()`` y, z, t, tdg and cy (engine.py:52-374); unitary and kraus
ops whose matrices are not in the circuit's caches are skipped.  Pass
``reference_dispatch=False`` to apply those gates instead.
"""
from __future__ import annotations

import math
import threading
from collections import OrderedDict
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from . import gates as G
from .numerics import get_backend
from .pauli import PauliSum

_STATE_SKIP = {"y", "z", "t", "tdg", "cy", "cry"}
_RUN_SKIP = {"y", "z", "t", "tdg", "cy"}
# ops after which run() attenuates <Z> of their wires under depolarizing noise
_ATTEN_1W = {"h", "rz", "rx", "ry", "x", "s", "sdg"}
_ATTEN_2W = {"cx", "cry", "cz", "iswap", "swap", "rxx", "ryy", "rzz"}


def _torch():
    import torch

    return torch


class _Lowerer:
    """Builds tqsv_op rows and segments for one circuit."""

    def __init__(self, n: int, track_grad: bool):
        self.n = n
        self.rows: List[tuple] = []
        self.segments: List[tuple] = []
        self.params: List[Any] = []  # torch tensors requiring grad, by param index
        self.track_grad = track_grad
        self.structure: List[tuple] = []

    def bit(self, q: int) -> int:
        return self.n - 1 - int(q)

    def angle(self, theta: Any) -> Tuple[float, int]:
        """(value, param index or -1)."""
        if hasattr(theta, "detach"):
            val = float(theta.item()) if theta.numel() == 1 else float(theta.detach().reshape(-1)[0])
            if self.track_grad and getattr(theta, "requires_grad", False):
                self.params.append(theta)
                return val, len(self.params) - 1
            return val, -1
        if isinstance(theta, str):
            raise TypeError(f"unbound symbolic parameter {theta!r}")
        return float(theta), -1

    @staticmethod
    def _m(mat: Any) -> List[float]:
        if isinstance(mat, list):  # already the 32 row floats
            return mat
        flat = np.asarray(mat, dtype=np.complex128).reshape(-1)
        out = np.zeros(32)
        out[0: 2 * flat.size: 2] = flat.real
        out[1: 2 * flat.size: 2] = flat.imag
        return list(out)

    def mat1(self, q: int, mat: np.ndarray, *, cmask: int = 0, cval: int = 0, param: int = -1,
             gx: int = 0, gz: int = 0) -> None:
        b = self.bit(q)
        self.rows.append((nat.OP_MAT1, (b, -1), param, cmask, cval, 0, 0, 0.0, gx, gz,
                          1.0 if param >= 0 else 0.0, self._m(mat)))

    def mat2(self, q0: int, q1: int, mat: np.ndarray, *, param: int = -1, gx: int = 0,
             gz: int = 0) -> None:
        b0, b1 = self.bit(q0), self.bit(q1)
        self.rows.append((nat.OP_MAT2, (b0, b1), param, 0, 0, 0, 0, 0.0, gx, gz,
                          1.0 if param >= 0 else 0.0, self._m(mat)))

    def prot(self, xmask: int, zmask: int, theta: float, param: int = -1) -> None:
        self.rows.append((nat.OP_PROT, (-1, -1), param, 0, 0, xmask, zmask, theta, 0, 0,
                          1.0 if param >= 0 else 0.0, [0.0] * 32))

    def flush(self) -> None:
        if self.rows:
            arr = np.array(self.rows, dtype=nat.OP_DTYPE)
            self.segments.append(("ops", arr))
            self.rows = []


_FIXED_ROWS: Dict[str, List[float]] = {}


def _fixed_row(name: str) -> List[float]:
    """Row floats of an unparameterised gate matrix (built once)."""
    row = _FIXED_ROWS.get(name)
    if row is None:
        mat = {"h": G.gate_h, "x": G.gate_x, "y": G.gate_y, "cz": G.gate_cz_4x4}[name]()
        row = _FIXED_ROWS[name] = _Lowerer._m(mat)
    return list(row)


def _rot_row(name: str, theta: float) -> List[float]:
    """Row floats of RX / RY / RZ(theta) (gates.py: cos(t/2) I - i sin(t/2) P,
    RY = [[c, -s], [s, c]]) without forming the matrix."""
    t = float(theta)
    c, s = math.cos(0.5 * t), math.sin(0.5 * t)
    if name == "rx":
        head = [c, 0.0, 0.0, -s, 0.0, -s, c, 0.0]
    elif name == "ry":
        head = [c, 0.0, -s, 0.0, s, 0.0, c, 0.0]
    else:
        head = [c, -s, 0.0, 0.0, 0.0, 0.0, c, s]
    return head + [0.0] * 24


def lower(circuit: Any, *, mode: str = "state", reference_dispatch: bool = True,
          track_grad: bool = False, noise: Optional[Dict[str, Any]] = None):
    """Lower circuit.ops to segments.  Returns (lowerer, measures, z_atten)."""
    n = int(circuit.num_qubits)
    L = _Lowerer(n, track_grad)
    skip = (_STATE_SKIP if mode == "state" else _RUN_SKIP) if reference_dispatch else set()
    ucache = getattr(circuit, "_unitary_cache", {}) or {}
    kcache = getattr(circuit, "_kraus_cache", {}) or {}
    measures: List[int] = []
    z_atten = [1.0] * n if noise is not None else None
    nfac = 1.0
    if noise is not None and str(noise.get("type", "")).lower() == "depolarizing":
        nfac = max(0.0, 1.0 - 4.0 * float(noise.get("p", 0.0)) / 3.0)

    def atten(*qs: int) -> None:
        if z_atten is not None:
            for q in qs:
                z_atten[int(q)] *= nfac

    for op in circuit.ops:
        if not isinstance(op, (list, tuple)) or not op:
            continue
        name = op[0]
        if name in skip:
            continue
        if name == "h":
            L.mat1(op[1], _fixed_row("h"))
        elif name in ("rx", "ry", "rz"):
            q = int(op[1])
            th, p = L.angle(op[2])
            b = 1 << L.bit(q)
            if name == "rx":
                L.mat1(q, _rot_row(name, th), param=p, gx=b)
            elif name == "ry":
                L.mat1(q, _rot_row(name, th), param=p, gx=b, gz=b)
            else:
                L.mat1(q, _rot_row(name, th), param=p, gz=b)
        elif name in ("cx", "cy"):
            c, t = int(op[1]), int(op[2])
            cm = 1 << L.bit(c)
            L.mat1(t, _fixed_row("x" if name == "cx" else "y"), cmask=cm, cval=cm)
        elif name == "cry":
            c, t = int(op[1]), int(op[2])
            th, p = L.angle(op[3])
            if p >= 0:
                raise NotImplementedError("gradients of cry are not supported")
            cm = 1 << L.bit(c)
            L.mat1(t, G.gate_ry(th), cmask=cm, cval=cm)
        elif name == "cz":
            L.mat2(op[1], op[2], _fixed_row("cz"))
        elif name == "iswap":
            L.mat2(op[1], op[2], G.gate_iswap_4x4())
        elif name == "swap":
            L.mat2(op[1], op[2], G.gate_swap_4x4())
        elif name in ("rxx", "ryy", "rzz"):
            a, bq = int(op[1]), int(op[2])
            th, p = L.angle(op[3])
            m = (1 << L.bit(a)) | (1 << L.bit(bq))
            if name == "rxx":
                L.mat2(a, bq, G.gate_rxx(th), param=p, gx=m)
            elif name == "ryy":
                L.mat2(a, bq, G.gate_ryy(th), param=p, gx=m, gz=m)
            else:
                L.mat2(a, bq, G.gate_rzz(th), param=p, gz=m)
        elif name == "x":
            L.mat1(op[1], G.gate_x())
        elif name == "y":
            L.mat1(op[1], G.gate_y())
        elif name == "z":
            L.mat1(op[1], G.gate_z())
        elif name == "s":
            L.mat1(op[1], G.gate_s())
        elif name == "sdg":
            L.mat1(op[1], G.gate_sd())
        elif name == "t":
            L.mat1(op[1], G.gate_t())
        elif name == "tdg":
            L.mat1(op[1], G.gate_td())
        elif name == "pauli_rot":
            fac, th = op[1], op[2]
            xm = zm = 0
            for P, q in fac:
                b = 1 << L.bit(q)
                if P in ("X", "Y"):
                    xm |= b
                if P in ("Z", "Y"):
                    zm |= b
            val, p = L.angle(th)
            L.prot(xm, zm, val, p)
        elif name == "measure_z":
            measures.append(int(op[1]))
            continue
        elif name == "barrier":
            continue
        elif name in ("project_z", "reset"):
            L.flush()
            keep = int(op[2]) if name == "project_z" else 0
            L.segments.append(("project", L.bit(int(op[1])), keep))
            continue
        elif name == "unitary":
            key = str(op[-1])
            mat = ucache.get(key)
            if mat is None:
                continue
            mat = np.asarray(mat.detach().cpu().numpy() if hasattr(mat, "detach") else mat,
                             dtype=np.complex128)
            if len(op) == 3:
                L.mat1(op[1], mat.reshape(2, 2))
            elif len(op) == 4:
                L.mat2(op[1], op[2], mat.reshape(4, 4))
            else:
                continue
        elif name == "kraus":
            kops = kcache.get(str(op[2]))
            if kops is None:
                continue
            L.flush()
            status = float(op[3]) if len(op) > 3 else None
            L.segments.append(("kraus", L.bit(int(op[1])),
                               [np.asarray(k, dtype=np.complex128) for k in kops], status))
            continue
        elif name in ("pulse", "pulse_inline"):
            raise NotImplementedError("pulse-level ops are not part of the tyxonq_amd engine")
        else:
            continue  # unsupported ops are ignored, as in the reference (engine.py:372-374)
        if name in _ATTEN_1W:
            atten(op[1])
        elif name in _ATTEN_2W or name == "unitary":
            # the op's wires, never its angle (engine.py:56-141: _attenuate(.., [q]) / [c, t])
            atten(*op[1:3] if name != "unitary" else op[1:-1])
    L.flush()
    return L, measures, z_atten


# ---- plan / Hamiltonian caches ------------------------------------------------

class _LRU:
    def __init__(self, size: int):
        self.size = size
        self.d: "OrderedDict[Any, Any]" = OrderedDict()

    def get(self, key: Any) -> Any:
        v = self.d.get(key)
        if v is not None:
            self.d.move_to_end(key)
        return v

    def put(self, key: Any, val: Any) -> None:
        self.d[key] = val
        self.d.move_to_end(key)
        while len(self.d) > self.size:
            self.d.popitem(last=False)


_PLANS = _LRU(64)
_HAMS = _LRU(32)


def _structure_key(n: int, ops: np.ndarray, opts: tuple) -> tuple:
    return (n, opts, ops["kind"].tobytes(), ops["t"].tobytes(), ops["cmask"].tobytes(),
            ops["cval"].tobytes(), ops["xmask"].tobytes(), ops["zmask"].tobytes(),
            ops["param"].tobytes())


def get_plan(n: int, ops: np.ndarray, *, tile_bits: int = 0, max_cost: int = 0,
             fuse: bool = True, min_qubits: int = 0) -> nat.Plan:
    """A plan for ops, reusing a cached plan of the same structure."""
    opts = (tile_bits, max_cost, fuse, min_qubits)
    key = _structure_key(n, ops, opts)
    plan = _PLANS.get(key)
    if plan is None:
        plan = nat.Plan(n, ops, tile_bits=tile_bits, max_cost=max_cost, fuse=fuse,
                        min_qubits=min_qubits)
        _PLANS.put(key, plan)
    else:
        plan.update(ops)
    return plan


def get_ham(n: int, obs: Any) -> Tuple[nat.Ham, PauliSum]:
    ps = PauliSum.coerce(obs, n)
    x, z, c = ps.to_masks()
    key = (n, x.tobytes(), z.tobytes(), c.tobytes())
    h = _HAMS.get(key)
    if h is None:
        h = nat.Ham(n, x, z, c)
        _HAMS.put(key, h)
    return h, ps


# ---- autograd ---------------------------------------------------------------

def _state_function():
    torch = _torch()

    class _CircuitState(torch.autograd.Function):
        """psi(theta) = U(theta)|psi0>; backward by the adjoint (uncompute) sweep."""

        @staticmethod
        def forward(ctx, plan, psi0, *params):
            psi = psi0.clone()
            plan.run(psi)
            ctx.plan = plan
            # The plan object is shared through the structure cache: a later
            # circuit of the same structure re-parameterises it in place.  Keep
            # this circuit's op rows, and restore them before the reverse sweep.
            ctx.ops = plan.ops
            ctx.meta = [(p.shape, p.dtype, p.device) for p in params]
            ctx.save_for_backward(psi)
            return psi

        @staticmethod
        def backward(ctx, gpsi):
            (psi,) = ctx.saved_tensors
            lam = gpsi.to(torch.complex64).contiguous().clone()
            work = psi.detach().clone()
            if ctx.plan.ops is not ctx.ops:
                ctx.plan.update(ctx.ops)
            g = ctx.plan.vjp(work, lam, len(ctx.meta))
            return (None, None) + _grads_like(g, ctx.meta, 1.0)

    return _CircuitState


def _energy_function():
    torch = _torch()

    class _Energy(torch.autograd.Function):
        """E(theta) = <psi(theta)|H|psi(theta)> with the gradient from one
        forward and one reverse sweep (computed eagerly in forward)."""

        @staticmethod
        def forward(ctx, plan, ham, psi0, *params):
            psi = psi0.clone()
            lam = torch.empty_like(psi)
            e, g = plan.adjoint(ham, psi, lam, len(params))
            ctx.grad = g
            ctx.meta = [(p.shape, p.dtype, p.device) for p in params]
            return torch.tensor(e, dtype=torch.float64)

        @staticmethod
        def backward(ctx, ge):
            return (None, None, None) + _grads_like(ctx.grad, ctx.meta, float(ge))

    return _Energy


def _product_function():
    torch = _torch()

    class _ProductExpectation(torch.autograd.Function):
        """Re <O psi|psi> for a product O of 1- and 2-qubit factors."""

        @staticmethod
        def forward(ctx, psi, factors):
            phi = psi.detach().clone()
            _apply_factors(phi, factors, dagger=False)
            val = nat.dot(phi, psi.detach()).real
            ctx.factors = factors
            ctx.save_for_backward(psi)
            return torch.tensor(val, dtype=torch.float64)

        @staticmethod
        def backward(ctx, ge):
            (psi,) = ctx.saved_tensors
            a = psi.detach().clone()
            b = psi.detach().clone()
            _apply_factors(a, ctx.factors, dagger=False)
            _apply_factors(b, ctx.factors, dagger=True)
            return (a + b) * float(ge), None

    return _ProductExpectation


def _grads_like(g: np.ndarray, meta, scale: float) -> tuple:
    torch = _torch()
    return tuple(torch.full(shape, scale * float(v), dtype=dtype, device=dev)
                 for v, (shape, dtype, dev) in zip(g, meta))


def _apply_factors(psi, factors, dagger: bool) -> None:
    seq = list(factors)
    if dagger:
        seq = [(np.conj(m).T, bits) for m, bits in reversed(seq)]
    for m, bits in seq:
        if len(bits) == 1:
            nat.apply_1q(psi, bits[0], m)
        else:
            nat.apply_2q(psi, bits[0], bits[1], m)


def bitstrings(index: Any, n: int) -> List[str]:
    """Big-endian n-bit strings of basis indices (qubit 0 first), built as
    one byte array instead of a format call per index."""
    idx = np.asarray(index, dtype=np.uint64).reshape(-1)
    if idx.size == 0:
        return []
    bits = np.unpackbits(idx.astype(">u8").view(np.uint8).reshape(-1, 8), axis=1)[:, 64 - n:]
    blob = np.ascontiguousarray(bits + 48).tobytes().decode("ascii")
    return [blob[i:i + n] for i in range(0, len(blob), n)]


_NO_SHARE = {"kraus", "project_z", "reset", "pulse", "pulse_inline"}


def _op_key(op: Any) -> Any:
    """A hashable, comparable form of an op tuple (tensor and array angles by
    value); None for ops that must not be shared between circuits."""
    if not isinstance(op, (list, tuple)) or not op or op[0] in _NO_SHARE:
        return None
    key = []
    for v in op:
        if hasattr(v, "detach"):
            a = v.detach().cpu().numpy()
            key.append(("t", a.shape, a.tobytes()))
        elif isinstance(v, np.ndarray):
            key.append(("a", v.shape, v.tobytes()))
        elif isinstance(v, (list, tuple)):
            key.append(("l", repr(v)))
        else:
            key.append(v)
    return tuple(key)


def _op_keys(circuit: Any) -> List[Any]:
    return [_op_key(op) for op in circuit.ops]


def _common_prefix(a: Sequence[Any], b: Sequence[Any]) -> int:
    k = 0
    for x, y in zip(a, b):
        if x is None or y is None or x != y:
            break
        k += 1
    return k


def _with_ops(circuit: Any, ops: List[Any]) -> Any:
    from .circuit import Circuit

    c = Circuit(int(circuit.num_qubits), ops=list(ops))
    for attr in ("_unitary_cache", "_kraus_cache"):
        if hasattr(circuit, attr):
            setattr(c, attr, getattr(circuit, attr))
    return c


# ---- engine -----------------------------------------------------------------

class StatevectorEngine:
    name = "statevector"
    capabilities = {"supports_shots": True}

    def __init__(self, backend_name: Optional[str] = None, *, reference_dispatch: bool = True,
                 fuse: bool = True, tile_bits: int = 0, max_cost: int = 0):
        self.backend = get_backend(backend_name)
        self.reference_dispatch = reference_dispatch
        self.fuse = fuse
        self.tile_bits = tile_bits
        self.max_cost = max_cost
        self.last_state = None  # final state of the last run() (reused by the driver)
        self._batch = None  # single-launch adjoint for small states (created on first use)
        self._batch_lock = threading.Lock()

    # -- core --
    def _init_state(self, circuit: Any, use_inputs: bool):
        torch = _torch()
        n = int(circuit.num_qubits)
        if n < 1:
            raise ValueError("the GPU engine needs at least one qubit")
        init = getattr(circuit, "_initial_state", None) if use_inputs else None
        if init is None:
            return nat.new_state(n, 0)
        arr = init.detach() if hasattr(init, "detach") else torch.as_tensor(np.asarray(init))
        arr = arr.reshape(-1)
        if arr.numel() != (1 << n):
            raise ValueError(f"initial state has {arr.numel()} amplitudes, expected {1 << n}")
        return arr.to(device=nat.device(), dtype=torch.complex64).contiguous().clone()

    def _plan(self, n: int, ops: np.ndarray) -> nat.Plan:
        return get_plan(n, ops, tile_bits=self.tile_bits, max_cost=self.max_cost, fuse=self.fuse)

    def _execute(self, L: _Lowerer, psi) -> None:
        for seg in L.segments:
            kind = seg[0]
            if kind == "ops":
                self._plan(L.n, seg[1]).run(psi)
            elif kind == "project":
                nat.project_z(psi, seg[1], seg[2])
            elif kind == "kraus":
                self._kraus(psi, seg[1], seg[2], seg[3])

    @staticmethod
    def _kraus(psi, bit: int, kops: Sequence[np.ndarray], status: Optional[float]) -> None:
        """Stochastic unravelling (kernels/statevector.py:132-218)."""
        probs = []
        for K in kops:
            tmp = psi.clone()
            nat.apply_1q(tmp, bit, K)
            probs.append(nat.norm2(tmp))
        p = np.asarray(probs, dtype=np.float64)
        p = p / p.sum()
        if status is None:
            import random

            status = random.random()
        sel = 0
        for i, c in enumerate(np.cumsum(p)):
            if status <= float(c):
                sel = i
                break
        nat.apply_1q(psi, bit, kops[sel])
        nrm = nat.norm2(psi)
        if nrm > 0:
            nat.scale(psi, 1.0 / math.sqrt(nrm))

    def _wants_grad(self) -> bool:
        if self.backend.name != "pytorch":
            return False
        return _torch().is_grad_enabled()

    def state_tensor(self, circuit: Any, *, mode: str = "state"):
        """Final state as a GPU complex64 tensor (differentiable with the
        pytorch backend when angles require grad)."""
        L, _, _ = lower(circuit, mode=mode, reference_dispatch=self.reference_dispatch,
                        track_grad=self._wants_grad())
        psi0 = self._init_state(circuit, use_inputs=(mode == "state"))
        if L.params:
            if len(L.segments) != 1 or L.segments[0][0] != "ops":
                raise NotImplementedError("gradients through measurement-type ops are not supported")
            plan = self._plan(L.n, L.segments[0][1])
            return _state_function().apply(plan, psi0, *L.params)
        self._execute(L, psi0)
        return psi0

    # -- engine contract --
    def state(self, circuit: Any) -> Any:
        """engine.py:897-1039."""
        psi = self.state_tensor(circuit, mode="state")
        if self.backend.name == "numpy":
            return psi.detach().cpu().numpy().astype(np.complex128)
        return psi

    def run(self, circuit: Any, shots: Optional[int] = None, **kwargs: Any) -> Dict[str, Any]:
        """engine.py:43-473."""
        shots = int(shots or 0)
        n = int(circuit.num_qubits)
        use_noise = bool(kwargs.get("use_noise", False))
        noise = (kwargs.get("noise") or {}) if use_noise else None
        L, measures, z_atten = lower(circuit, mode="run", reference_dispatch=self.reference_dispatch,
                                     track_grad=False, noise=noise)
        psi = self._init_state(circuit, use_inputs=False)
        self._execute(L, psi)
        self.last_state = psi
        if shots > 0 and len(measures) > 0:
            counts = self._sample(psi, n, shots, kwargs)
            return {"result": counts,
                    "metadata": {"shots": shots, "backend": self.backend.name,
                                 "three_level": bool(kwargs.get("three_level", False))}}
        expectations: Dict[str, float] = {}
        for q in measures:
            val = nat.expect_z(psi, n - 1 - q)
            if z_atten is not None:
                val *= z_atten[q]
            expectations[f"Z{q}"] = float(val)
        return {"expectations": expectations,
                "metadata": {"shots": shots, "backend": self.backend.name}}

    def run_many(self, circuits: Sequence[Any], shots: Sequence[int], **kwargs: Any
                 ) -> List[Dict[str, Any]]:
        """run() of several circuits with shots > 0 that simulates the ops
        consecutive circuits have in common once.  Batches such as a base
        circuit followed by per-group basis rotations (hea_device_runtime.py:
        150-175) then cost one simulation plus one short suffix per circuit.
        Circuit i keeps a copy of its state after the ops it shares with
        circuit i+1 (same width, equal leading ops without kraus or reset type
        ops), and circuit i+1 starts there.  Each circuit draws its counts as
        run() does; with ``seed`` circuit i uses the SeedSequence child
        (seed, i)."""
        shots = [int(s) for s in shots]
        if len(shots) != len(circuits):
            raise ValueError("shots list length must match number of circuits")
        if any(s <= 0 for s in shots):
            raise ValueError("run_many samples counts: every shot count must be positive")
        use_noise = bool(kwargs.get("use_noise", False))
        noise = (kwargs.get("noise") or {}) if use_noise else None
        keys = [_op_keys(c) for c in circuits]
        out: List[Dict[str, Any]] = []
        snap: Optional[Tuple[int, Any]] = None
        for i, c in enumerate(circuits):
            n = int(c.num_qubits)
            ops = list(c.ops)
            share = 0
            if i + 1 < len(circuits) and int(circuits[i + 1].num_qubits) == n:
                share = _common_prefix(keys[i], keys[i + 1])
            start, psi = 0, None
            if snap is not None and i > 0 and int(circuits[i - 1].num_qubits) == n \
                    and _common_prefix(keys[i - 1], keys[i]) == snap[0]:
                start, psi = snap
            snap = None
            if psi is None:
                psi = nat.new_state(n, 0)
            # measure ops of a prefix run by the previous circuit
            measures = [int(op[1]) for op in ops[:start]
                        if isinstance(op, (list, tuple)) and op and op[0] == "measure_z"]
            if share > 0 and share == start:
                snap = (share, psi.clone())
            cuts = [start] + ([share] if share > start else []) + [len(ops)]
            for a, b in zip(cuts[:-1], cuts[1:]):
                L, m, _ = lower(_with_ops(c, ops[a:b]), mode="run",
                                reference_dispatch=self.reference_dispatch, track_grad=False,
                                noise=noise)
                self._execute(L, psi)
                measures += m
                if b == share and share > start:
                    snap = (share, psi.clone())
            self.last_state = psi
            kw = dict(kwargs)
            if kwargs.get("seed") is not None:
                kw["seed"] = int(np.random.SeedSequence([int(kwargs["seed"]) & (2 ** 64 - 1), i])
                                 .generate_state(1, dtype=np.uint64)[0])
            meta = {"shots": shots[i], "backend": self.backend.name,
                    "three_level": bool(kwargs.get("three_level", False))}
            if measures:
                out.append({"result": self._sample(psi, n, shots[i], kw), "metadata": meta})
            else:
                out.append({"expectations": {}, "metadata": meta})
            del psi
        return out

    def _sample(self, psi, n: int, shots: int, kwargs: Dict[str, Any]) -> Dict[str, int]:
        """Counts over the computational basis (engine.py:377-465) as big-endian
        bitstrings.  Without noise the shots are drawn on the GPU (inverse CDF
        of |psi|^2, Philox stream keyed by ``seed``) and only the distinct
        states come back.  Readout or depolarizing mixing of the probabilities
        (engine.py:387-408) is applied to |psi|^2 on the host before a
        multinomial draw, as in the reference."""
        if not bool(kwargs.get("use_noise", False)):
            seed = kwargs.get("seed")
            if seed is None:
                seed = int(np.random.SeedSequence().generate_state(1, dtype=np.uint64)[0])
            try:
                idx, cnt = nat.sample_counts(psi, shots, int(seed))
                return dict(zip(bitstrings(idx, n), cnt.tolist()))
            except ValueError:
                if nat.norm2(psi) > 0.0:
                    raise
                # a zero state: the reference draws uniformly (engine.py:409-412)
        p = nat.probs(psi).cpu().numpy()
        dim = p.size
        if bool(kwargs.get("use_noise", False)):
            noise = kwargs.get("noise", {}) or {}
            ntype = str(noise.get("type", "")).lower()
            if ntype == "readout":
                cals = noise.get("cals", {}) or {}
                A = None
                for q in range(n):
                    m = np.asarray(cals.get(q, np.eye(2)), dtype=float)
                    A = m if A is None else np.kron(A, m)
                p = A @ p
            elif ntype == "depolarizing":
                alpha = max(0.0, min(1.0, 4.0 * float(noise.get("p", 0.0)) / 3.0))
                p = (1.0 - alpha) * p + alpha / dim
            p = np.clip(p, 0.0, 1.0)
            s = float(p.sum())
            p = p / (s if s > 1e-12 else 1.0)
        s = float(p.sum())
        p = p / s if s > 0 else np.full(dim, 1.0 / dim)
        rng = np.random.default_rng(kwargs.get("seed"))
        idx = rng.choice(dim, size=shots, p=p)
        cnt = np.bincount(idx, minlength=dim)
        out: Dict[str, int] = {}
        for i in np.nonzero(cnt)[0]:
            out[format(int(i), f"0{n}b")] = int(cnt[i])
        return out

    def expval(self, circuit: Any, obs: Any, **kwargs: Any) -> float:
        """Analytic <H> for a Pauli sum (engine.py:475-484 takes an OpenFermion
        QubitOperator; its ``terms`` dict is accepted here as well)."""
        e = self.energy(circuit, obs)
        return float(e.detach()) if hasattr(e, "detach") else float(e)

    def energy(self, circuit: Any, obs: Any) -> Any:
        """<psi|H|psi> of a Pauli sum in fused FP64 reductions; differentiable
        (adjoint sweep) with the pytorch backend."""
        n = int(circuit.num_qubits)
        ham, _ = get_ham(n, obs)
        L, _, _ = lower(circuit, mode="state", reference_dispatch=self.reference_dispatch,
                        track_grad=self._wants_grad())
        psi0 = self._init_state(circuit, use_inputs=True)
        if L.params:
            if len(L.segments) != 1 or L.segments[0][0] != "ops":
                raise NotImplementedError("gradients through measurement-type ops are not supported")
            plan = self._plan(n, L.segments[0][1])
            return _energy_function().apply(plan, ham, psi0, *L.params)
        self._execute(L, psi0)
        e, _ = ham.expval(psi0)
        if self.backend.name == "pytorch":
            return _torch().tensor(e, dtype=_torch().float64)
        return np.float64(e)

    def energy_and_grad(self, circuit: Any, obs: Any) -> Tuple[float, np.ndarray]:
        """(E, dE/dparam) where the parameters are the circuit's tensor-valued
        angles that require grad, in op order (one forward + one reverse sweep)."""
        torch = _torch()
        n = int(circuit.num_qubits)
        ham, _ = get_ham(n, obs)
        L, _, _ = lower(circuit, mode="state", reference_dispatch=self.reference_dispatch,
                        track_grad=True)
        if len(L.segments) != 1 or L.segments[0][0] != "ops":
            raise NotImplementedError("energy_and_grad needs a purely unitary circuit")
        psi = self._init_state(circuit, use_inputs=True)
        if n <= nat.BATCH_QUBITS and not self.tile_bits:
            # the whole sweep in one launch, state and lambda in LDS
            plan = get_plan(n, L.segments[0][1], max_cost=self.max_cost, fuse=self.fuse,
                            min_qubits=nat.BATCH_QUBITS)
            with self._batch_lock:
                if self._batch is None:
                    self._batch = nat.Batch()
                ((e, g),) = self._batch.adjoint([plan], [ham], [psi], [len(L.params)])
            return e, g
        plan = self._plan(n, L.segments[0][1])
        lam = torch.empty_like(psi)
        return plan.adjoint(ham, psi, lam, len(L.params))

    def expectation_product(self, circuit: Any, pauli_ops: Sequence[Any]) -> Any:
        """circuit.py:1751-1796: <psi|O|psi> for (matrix, qubits) factors."""
        n = int(circuit.num_qubits)
        factors = []
        for spec in pauli_ops:
            if not isinstance(spec, (tuple, list)) or len(spec) != 2:
                raise ValueError(f"Each Pauli operator must be (gate, qubits), got {spec}")
            gate, qubits = spec
            if hasattr(gate, "tensor"):
                gate = gate.tensor
            elif callable(gate):
                gate = gate()
            if hasattr(gate, "detach"):
                gate = gate.detach().cpu().numpy()
            m = np.asarray(gate, dtype=np.complex128)
            if not isinstance(qubits, (list, tuple)):
                qubits = [qubits]
            if len(qubits) == 1:
                factors.append((m.reshape(2, 2), [n - 1 - int(qubits[0])]))
            elif len(qubits) == 2:
                factors.append((m.reshape(4, 4), [n - 1 - int(qubits[0]), n - 1 - int(qubits[1])]))
            else:
                raise NotImplementedError(f"Pauli products on {len(qubits)} qubits not yet supported")
        psi = self.state_tensor(circuit, mode="state")
        if self.backend.name == "pytorch":
            return _product_function().apply(psi, factors)
        phi = psi.detach().clone()
        _apply_factors(phi, factors, dagger=False)
        return np.float64(nat.dot(phi, psi.detach()).real)

    def probability(self, circuit: Any) -> np.ndarray:
        """engine.py:1041-1047."""
        psi = self.state_tensor(circuit, mode="state").detach()
        return nat.probs(psi).cpu().numpy()

    def amplitude(self, circuit: Any, bitstring: str) -> complex:
        """engine.py:1049-1059 (big-endian: q0 is the leftmost character)."""
        n = int(circuit.num_qubits)
        if len(bitstring) != n:
            raise ValueError("bitstring length must equal num_qubits")
        idx = 0
        for ch in bitstring:
            idx = (idx << 1) | (1 if ch == "1" else 0)
        psi = self.state_tensor(circuit, mode="state").detach()
        return complex(psi[idx].cpu().item())

    def perfect_sampling(self, circuit: Any, *, rng: Optional[np.random.Generator] = None):
        """engine.py:1061-1072."""
        n = int(circuit.num_qubits)
        p = self.probability(circuit)
        rng = np.random.default_rng() if rng is None else rng
        idx = int(rng.choice(1 << n, p=p / p.sum()))
        return format(idx, f"0{n}b"), float(p[idx])
