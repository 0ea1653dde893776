"""Device routing for the simulator provider.

Restates the simulator half of devices/base.py and
devices/simulators/driver.py of the reference:
  * ``run(provider, device, circuit=..., source=..., shots=..., **opts)``
    returns a list of ``DeviceTask`` (base.py:252-393);
  * the driver selects the engine by device name (driver.py:20-30), runs
    each circuit and wraps the result in the schema
    {result, expectations, probabilities, statevector, result_meta,
    uni_status, error} (driver.py:86-140);
  * ``DeviceTask.get_result`` normalises it the way base.py:20-36 and
    base.py:445-488 do.
With shots == 0 the reference simulates the circuit twice to attach the
probabilities and the statevector (driver.py:115-122); here the state of the
single run is reused.  Only provider "simulator"/"local" and the
statevector device are served: hardware and cloud providers are outside
this engine.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Dict, List, Optional, Sequence, Union
from uuid import uuid4

import numpy as np

_NOISE = {"enabled": False, "config": None}
_DEFAULT_NOISE = {"type": "depolarizing", "p": 0.0}


def enable_noise(enabled: bool = True, config: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
    _NOISE["enabled"] = bool(enabled)
    if config is not None:
        _NOISE["config"] = dict(config)
    return {"enabled": _NOISE["enabled"], "config": _NOISE["config"] or {}}


def is_noise_enabled() -> bool:
    return bool(_NOISE["enabled"])


def get_noise_config() -> Dict[str, Any]:
    return dict(_NOISE["config"] or {})


def _inject_noise(kwargs: Dict[str, Any]) -> Dict[str, Any]:
    use_noise = bool(kwargs.get("use_noise", _NOISE["enabled"]))
    if not use_noise:
        return kwargs
    new = dict(kwargs)
    new["use_noise"] = True
    if new.get("noise") is None:
        new["noise"] = _NOISE["config"] or dict(_DEFAULT_NOISE)
    return new


class SimTask:
    """An already-finished simulator task (driver.py:6-17)."""

    def __init__(self, id: str, device: str, result: Dict[str, Any]):
        self.id = id
        self.device = device
        self._result = result
        self.task_info = None
        self.async_result = False
        # final GPU state of a shots = 0 run, for the analytic post-processing
        # of Circuit.run; released by it
        self.device_state = None

    def get_result(self, *, wait: bool = False, poll_interval: float = 0.0,
                   timeout: float = 0.0) -> Dict[str, Any]:
        return self._result


@dataclass
class DeviceTask:
    provider: str
    device: str
    handle: Any
    async_result: bool

    def get_result(self, *, wait: bool = False, poll_interval: float = 2.0,
                   timeout: float = 60.0) -> Dict[str, Any]:
        info = self.handle.get_result()
        unified = {
            "result": info.get("result") or info.get("results") or {},
            "result_meta": info,
            "uni_status": str(info.get("uni_status", "unknown")).lower(),
            "error": info.get("error", None),
        }
        result = dict(unified)
        result.setdefault("metadata", result["result_meta"])
        result.setdefault("expectations", {})
        result.setdefault("probabilities", None)
        return result

    def cancel(self) -> Dict[str, str]:
        return {"state": "cancelled"}


def _device_name(device: Optional[str]) -> str:
    d = "statevector" if device is None else str(device)
    return d.split("::")[-1] if "::" in d else d


def _select_engine(device: str):
    name = _device_name(device)
    if name in ("simulator:statevector", "statevector"):
        from .engine import StatevectorEngine

        return StatevectorEngine
    if name in ("simulator:density_matrix", "density_matrix"):
        from .density import DensityMatrixEngine

        return DensityMatrixEngine
    if name in ("simulator:mps", "mps", "matrix_product_state"):
        raise NotImplementedError(f"simulator device '{device}' is not part of tyxonq_amd")
    raise ValueError(f"Unsupported simulator device: {device}")


def list_devices(token: Optional[str] = None, **kws: Any) -> List[str]:
    return ["simulator::statevector", "simulator::density_matrix"]


def _state_dispatch_matches_run(eng: Any, c: Any) -> bool:
    """True when state() would apply exactly the ops run() applied, to the
    same start state: no inputs (run() starts from |0..0>, state() honours
    circuit._initial_state, engine.py:46 vs 906-912), no cry under the
    reference dispatch (state() skips it, run() applies it, engine.py:914-1038
    vs 76-79) and no Kraus op without a status (each call draws its own
    trajectory)."""
    if getattr(c, "_initial_state", None) is not None:
        return False
    ref = bool(getattr(eng, "reference_dispatch", True))
    for op in getattr(c, "ops", []):
        if not isinstance(op, (list, tuple)) or not op:
            continue
        if op[0] == "cry" and ref:
            return False
        if op[0] == "kraus" and len(op) <= 3:
            return False
    return True


def _attached_state(eng: Any, c: Any) -> Any:
    """The state the driver attaches to a shots=0 result: eng.state(c)
    (driver.py:115-122).  The state of the run() just made is reused when the
    two dispatches provably coincide, which saves the reference's second
    simulation; otherwise state() is simulated.  Engines without state()
    raise, as in the reference, and the driver reports it in ``error``."""
    if not hasattr(eng, "state_tensor"):
        return eng.state(c)
    last = getattr(eng, "last_state", None)
    if last is not None and _state_dispatch_matches_run(eng, c):
        return last.detach()
    return eng.state_tensor(c, mode="state").detach()


def _run_one(eng: Any, device: str, c: Any, shots: int, opts: Dict[str, Any]) -> SimTask:
    error = ""
    out: Dict[str, Any] = {}
    try:
        out = eng.run(c, shots=shots, **opts)
    except Exception as e:  # the driver reports engine failures in the payload
        error = str(e)
    counts = out.get("result") or {}
    expectations = out.get("expectations") or {}
    meta = dict(out.get("metadata", {}))
    prob = statevec = None
    psi = None
    if int(shots) == 0 and not error:
        try:
            psi = _attached_state(eng, c)
            statevec = (psi.cpu().numpy() if hasattr(psi, "cpu") else np.asarray(psi)
                        ).astype(np.complex128)
            prob = np.abs(statevec) ** 2
            meta.setdefault("num_qubits", int(getattr(c, "num_qubits", 0)))
        except Exception as e:
            error = str(e)
    result = {"result": counts, "expectations": expectations, "probabilities": prob,
              "statevector": statevec, "result_meta": meta, "uni_status": "completed",
              "error": error}
    task = SimTask(id=str(uuid4()), device=device, result=result)
    if psi is not None and hasattr(psi, "is_cuda") and not error:
        task.device_state = psi
    return task


def run(*, provider: Optional[str] = None, device: Optional[str] = None,
        circuit: Optional[Any] = None, source: Optional[Any] = None,
        shots: Union[int, Sequence[int]] = 1024, **opts: Any) -> List[DeviceTask]:
    prov = provider or "simulator"
    dev = device or "statevector"
    if prov not in ("simulator", "local"):
        raise ValueError(f"provider '{prov}' is not served by tyxonq_amd (simulator only)")
    if source is not None and isinstance(source, str):
        raise ValueError("QASM sources are not accepted by tyxonq_amd; pass a Circuit")
    target = source if source is not None else circuit
    if target is None:
        raise ValueError("run requires either circuit or source")
    circuits = list(target) if isinstance(target, (list, tuple)) else [target]
    Engine = _select_engine(dev)
    eng = Engine()
    if isinstance(shots, (list, tuple)):
        shots_list = [int(s) for s in shots]
        if len(shots_list) != len(circuits):
            raise ValueError("shots list length must match number of circuits")
    else:
        shots_list = [int(shots)] * len(circuits)
    if len(circuits) == 1 and shots_list[0] == 0 and "observable" in opts:
        e = eng.expval(circuits[0], opts["observable"])
        task = SimTask(id="expval", device=dev, result={
            "result": {}, "expectations": {"expval": float(e)}, "probabilities": None,
            "statevector": None,
            "metadata": {"shots": 0, "backend": "analytic", "provider": prov, "device": dev}})
        return [DeviceTask(provider=prov, device=dev, handle=task, async_result=False)]
    kw = _inject_noise({k: v for k, v in opts.items() if k != "observable"})
    out = []
    if len(circuits) > 1 and all(s > 0 for s in shots_list) and hasattr(eng, "run_many"):
        # counts of a batch: leading ops shared by consecutive circuits run once
        try:
            outs = eng.run_many(circuits, shots_list, **kw)
        except Exception:
            outs = None
        if outs is not None:
            for c, s, o in zip(circuits, shots_list, outs):
                res = {"result": o.get("result") or {}, "expectations": o.get("expectations") or {},
                       "probabilities": None, "statevector": None,
                       "result_meta": dict(o.get("metadata", {})), "uni_status": "completed",
                       "error": ""}
                out.append(DeviceTask(provider=prov, device=dev,
                                      handle=SimTask(id=str(uuid4()), device=dev, result=res),
                                      async_result=False))
            return out
    for c, s in zip(circuits, shots_list):
        out.append(DeviceTask(provider=prov, device=dev, handle=_run_one(eng, dev, c, s, kw),
                              async_result=False))
    return out


def expval(*, provider: Optional[str] = None, device: Optional[str] = None,
           circuit: Any = None, observable: Any = None, **opts: Any) -> float:
    """Analytic expectation for simulator providers (base.py:396-417)."""
    prov = provider or "simulator"
    if prov not in ("simulator", "local"):
        raise ValueError("expval is supported only for simulator/local providers")
    if circuit is None or observable is None:
        raise ValueError("expval requires both circuit and observable")
    return float(_select_engine(device or "statevector")().expval(circuit, observable, **opts))
