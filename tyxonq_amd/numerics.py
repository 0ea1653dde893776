"""Numerics backend selection (the reference's ``tq.set_backend`` surface).

The reference selects a global ArrayBackend by name ("numpy", "pytorch",
"cupynumeric"), numerics/context.py:18-50 and numerics/api.py:179-252.  In
this engine the statevector always lives on the GPU as a torch complex64
tensor; the backend only decides what ``Circuit.state()`` hands back:
  * "numpy"   -> a host numpy complex128 array (no autograd), as the
                 reference's numpy backend returns;
  * "pytorch" -> the device tensor itself, differentiable through the
                 adjoint engine when circuit parameters require grad.
"""
from __future__ import annotations

import threading
from typing import Any, Callable, Optional

import numpy as np

_ALIASES = {
    "numpy": "numpy",
    "np": "numpy",
    "pytorch": "pytorch",
    "torch": "pytorch",
    "pt": "pytorch",
}

_state = threading.local()
_global = {"name": "numpy"}


class Backend:
    """Minimal backend handle: a name plus dtype constants."""

    def __init__(self, name: str):
        self.name = name
        self.complex128 = np.complex128
        self.complex64 = np.complex64
        self.float64 = np.float64
        self.float32 = np.float32

    def __repr__(self) -> str:
        return f"Backend({self.name!r})"


def normalize_backend_name(name: Any) -> str:
    if isinstance(name, Backend):
        return name.name
    key = str(name).lower()
    if key not in _ALIASES:
        raise ValueError(f"Unknown backend '{name}'. Supported: numpy, pytorch")
    return _ALIASES[key]


def set_backend(name_or_instance: Any) -> Backend:
    _global["name"] = normalize_backend_name(name_or_instance)
    return get_backend()


def get_backend(name: Optional[Any] = None) -> Backend:
    if name is None:
        name = getattr(_state, "name", None) or _global["name"]
    return Backend(normalize_backend_name(name))


class use_backend:
    """Context manager that selects a backend for the current thread."""

    def __init__(self, name: Any):
        self.name = normalize_backend_name(name)
        self.prev: Optional[str] = None

    def __enter__(self) -> Backend:
        self.prev = getattr(_state, "name", None)
        _state.name = self.name
        return get_backend()

    def __exit__(self, *exc: Any) -> bool:
        _state.name = self.prev
        return False


_dtype = {"name": "complex64"}


def set_dtype(dtype_str: str) -> str:
    """Record the requested complex dtype.

    Device states are complex64 with FP64 reductions; "complex128" results
    are returned to the host as complex128 arrays.  (The reference records the
    dtype but its statevector engine ignores it, numerics/__init__.py:39-58
    vs. kernels/statevector.py:19-25.)"""
    d = str(dtype_str).lower()
    if d not in ("complex64", "complex128"):
        raise ValueError("dtype must be 'complex64' or 'complex128'")
    _dtype["name"] = d
    return d


def get_dtype() -> str:
    return _dtype["name"]


def value_and_grad(fn: Callable, argnums: int | tuple = 0) -> Callable:
    """(value, grad) of a scalar function of real parameters via torch autograd.

    Mirrors PyTorchBackend.value_and_grad (numerics/backends/pytorch_backend.py:446-507):
    arguments at ``argnums`` become float64 leaf tensors; the function value
    and gradients are returned as numpy.  Inside ``fn`` the circuit
    expectation is differentiated by the adjoint engine instead of storing
    every intermediate state."""
    import torch

    idx = (argnums,) if isinstance(argnums, int) else tuple(argnums)

    def wrapped(*args: Any, **kwargs: Any):
        args = list(args)
        leaves = []
        for i in idx:
            x = args[i]
            t = (x.detach().clone() if isinstance(x, torch.Tensor)
                 else torch.tensor(np.asarray(x, dtype=np.float64)))
            t = t.to(torch.float64).requires_grad_(True)
            args[i] = t
            leaves.append(t)
        y = fn(*args, **kwargs)
        if not isinstance(y, torch.Tensor):
            y = torch.tensor(float(y), dtype=torch.float64)
        grads = torch.autograd.grad(y, leaves, allow_unused=True)
        grads = [np.zeros(tuple(l.shape)) if g is None else g.detach().cpu().numpy()
                 for g, l in zip(grads, leaves)]
        val = y.detach().cpu().numpy()
        val = val.item() if val.size == 1 else val
        return val, (grads[0] if len(grads) == 1 else tuple(grads))

    return wrapped
