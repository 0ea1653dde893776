"""tyxonq_amd - MI355X (gfx950) statevector engine behind TyxonQ's simulator path.

Drop-in surface (the parts of ``import tyxonq as tq`` on this path):

    import tyxonq_amd as tq
    tq.set_backend("pytorch")
    c = tq.Circuit(2).h(0).cx(0, 1)
    c.device(provider="simulator", device="statevector")
    psi = c.state()                       # GPU complex64 tensor
    e = c.expectation([(1.0, [("Z", 0), ("Z", 1)])])   # fused Pauli-sum energy
    res = c.run(shots=1024)               # counts, reference result schema

The numerics live in libtqsv.so (hand-written HIP for gfx950, C ABI in
include/tqsv.h); this package is the host-side mirror of the reference's
Circuit / engine / driver interfaces.
"""
from __future__ import annotations

from typing import Any, Dict, Optional

from . import gates
from .circuit import (
    Circuit,
    Hamiltonian,
    get_global_compile_defaults,
    get_global_device_defaults,
    get_global_postprocessing_defaults,
    set_global_compile_defaults,
    set_global_device_defaults,
    set_global_postprocessing_defaults,
)
from .devices import enable_noise, get_noise_config, is_noise_enabled
from .numerics import get_backend, get_dtype, set_backend, set_dtype, use_backend, value_and_grad
from .pauli import PauliSum

__version__ = "0.1.0"

__all__ = [
    "Circuit", "Hamiltonian", "PauliSum", "gates", "set_backend", "get_backend", "use_backend",
    "set_dtype", "get_dtype", "value_and_grad", "device", "get_device_defaults", "compile",
    "get_compile", "postprocessing", "get_postprocessing_defaults", "enable_noise",
    "is_noise_enabled", "get_noise_config", "run", "StatevectorEngine",
]


def device(**options: Any) -> Dict[str, Any]:
    return set_global_device_defaults(options)


def get_device_defaults() -> Dict[str, Any]:
    return get_global_device_defaults()


def compile(**options: Any) -> Dict[str, Any]:  # noqa: A001 - mirrors tq.compile
    return set_global_compile_defaults(options)


def get_compile() -> Dict[str, Any]:
    return get_global_compile_defaults()


def postprocessing(**options: Any) -> Dict[str, Any]:
    return set_global_postprocessing_defaults(options)


def get_postprocessing_defaults() -> Dict[str, Any]:
    return get_global_postprocessing_defaults()


def run(circuit: Any, *, provider: Optional[str] = None, device: Optional[str] = None,
        shots: int = 1024, wait: bool = True, **task_kws: Any) -> Any:
    """Submit circuit(s) to the simulator and return the result dict(s)."""
    from . import devices as _dev

    tasks = _dev.run(provider=provider, device=device, circuit=circuit, shots=shots, **task_kws)
    if not wait:
        return tasks
    res = [t.get_result(wait=True) for t in tasks]
    return res if isinstance(circuit, (list, tuple)) else res[0]


def __getattr__(name: str) -> Any:
    if name == "StatevectorEngine":
        from .engine import StatevectorEngine

        return StatevectorEngine
    raise AttributeError(name)
