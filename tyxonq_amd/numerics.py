"""Numerics backend selection (the reference's ``tq.set_backend`` surface).

The reference selects a global ArrayBackend by name ("numpy", "pytorch",
"cupynumeric"), numerics/context.py:18-50 and numerics/api.py:179-252.  In
this engine the statevector always lives on the GPU as a torch complex64
tensor; the backend decides what ``Circuit.state()`` hands back:
  * "numpy"   -> a host numpy complex128 array (no autograd), as the
                 reference's numpy backend returns;
  * "pytorch" -> the device tensor itself, differentiable through the
                 adjoint engine when circuit parameters require grad.
The backend object also carries the reference's ArrayBackend operations
(numerics/api.py:19-121: array creation, elementwise math, linear algebra,
random numbers, autodiff helpers) for user code around the simulator, as in
numerics/backends/numpy_backend.py and pytorch_backend.py.
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
_global: dict = {"name": "numpy", "inst": None}


class Backend:
    """A numerics backend: its name, dtype tokens and the array operations of
    the reference's ArrayBackend protocol (numerics/api.py:19-121), so that
    code written against ``tq.get_backend(...)`` (``nb.abs``, ``nb.einsum``,
    ``nb.square``, ...) runs unchanged.  The names and signatures are the
    protocol's; ``NumpyBackend`` and ``TorchBackend`` implement them.  The
    statevector itself never goes through these: it lives on the GPU and the
    engine's kernels act on it."""

    name = "base"
    xp: Any = None  # the array module with numpy-compatible names
    dtypestr = "complex128"  # default complex dtype of this instance
    rdtypestr = "float64"  # matching real dtype

    def __repr__(self) -> str:
        return f"Backend({self.name!r})"

    def set_dtype(self, dtype_str: str) -> tuple:
        """Default dtypes of this instance: (complex, real) constants
        (numerics/backends/numpy_backend.py:27-46)."""
        if dtype_str == "complex64":
            self.dtypestr, self.rdtypestr = "complex64", "float32"
            return self.complex64, self.float32
        if dtype_str == "complex128":
            self.dtypestr, self.rdtypestr = "complex128", "float64"
            return self.complex128, self.float64
        raise ValueError(f"Unsupported dtype: {dtype_str}. Use 'complex64' or 'complex128'.")

    # operations that numpy and torch spell alike
    def matmul(self, a: Any, b: Any) -> Any:
        return a @ b

    def einsum(self, subscripts: str, *operands: Any) -> Any:
        return self.xp.einsum(subscripts, *operands)

    def abs(self, a: Any) -> Any:
        return self.xp.abs(a)

    def real(self, a: Any) -> Any:
        return self.xp.real(a)

    def imag(self, a: Any) -> Any:
        return self.xp.imag(a)

    def conj(self, a: Any) -> Any:
        return self.xp.conj(a)

    def square(self, a: Any) -> Any:
        return self.xp.square(a)

    def kron(self, a: Any, b: Any) -> Any:
        return self.xp.kron(a, b)

    def exp(self, a: Any) -> Any:
        return self.xp.exp(a)

    def sin(self, a: Any) -> Any:
        return self.xp.sin(a)

    def cos(self, a: Any) -> Any:
        return self.xp.cos(a)

    def sqrt(self, a: Any) -> Any:
        return self.xp.sqrt(a)

    def log(self, a: Any) -> Any:
        return self.xp.log(a)

    def log2(self, a: Any) -> Any:
        return self.xp.log2(a)

    def sign(self, a: Any) -> Any:
        return self.xp.sign(a)

    def diag(self, a: Any) -> Any:
        return self.xp.diag(a)

    def zeros_like(self, a: Any) -> Any:
        return self.xp.zeros_like(a)

    def ones_like(self, a: Any) -> Any:
        return self.xp.ones_like(a)

    def outer(self, a: Any, b: Any) -> Any:
        return self.xp.outer(self.reshape(a, (-1,)), self.reshape(b, (-1,)))

    def dot(self, a: Any, b: Any) -> Any:
        return self.matmul(a, b)

    def allclose(self, a: Any, b: Any, rtol: float = 1e-5, atol: float = 1e-8) -> bool:
        return bool(np.allclose(self.to_numpy(a), self.to_numpy(b), rtol=rtol, atol=atol))

    def isclose(self, a: Any, b: Any, rtol: float = 1e-5, atol: float = 1e-8) -> Any:
        return np.isclose(self.to_numpy(a), self.to_numpy(b), rtol=rtol, atol=atol)

    def jit(self, fn: Callable) -> Callable:
        # kernels are compiled per plan by the engine; a user function runs as is
        return fn


class NumpyBackend(Backend):
    """The numpy implementation (host arrays, no autodiff)."""

    name = "numpy"
    xp = np
    complex64, complex128 = np.complex64, np.complex128
    float32, float64 = np.float32, np.float64
    int32, int64, int8 = np.int32, np.int64, np.int8
    bool = np.bool_
    int = np.int64

    def array(self, data: Any, dtype: Any = None) -> Any:
        return np.array(self.to_numpy(data), dtype=dtype)

    def asarray(self, data: Any, dtype: Any = None) -> Any:
        return np.asarray(self.to_numpy(data), dtype=dtype)

    def to_numpy(self, data: Any) -> np.ndarray:
        if hasattr(data, "detach"):  # a torch tensor handed to the numpy backend
            return data.detach().resolve_conj().resolve_neg().cpu().numpy()
        return np.asarray(data)

    def reshape(self, a: Any, shape: Any) -> Any:
        return np.reshape(a, shape)

    def moveaxis(self, a: Any, source: int, destination: int) -> Any:
        return np.moveaxis(a, source, destination)

    def sum(self, a: Any, axis: Any = None) -> Any:
        return np.sum(a, axis=axis)

    def mean(self, a: Any, axis: Any = None) -> Any:
        return np.mean(a, axis=axis)

    def zeros(self, shape: Any, dtype: Any = None) -> Any:
        return np.zeros(shape, dtype=dtype if dtype is not None else np.float64)

    def ones(self, shape: Any, dtype: Any = None) -> Any:
        return np.ones(shape, dtype=dtype if dtype is not None else np.float64)

    def eye(self, n: int, dtype: Any = None) -> Any:
        return np.eye(n, dtype=dtype if dtype is not None else np.float64)

    def stack(self, arrays: Any, axis: int = 0) -> Any:
        return np.stack(arrays, axis=axis)

    def concatenate(self, arrays: Any, axis: int = 0) -> Any:
        return np.concatenate(arrays, axis=axis)

    def arange(self, start: Any, stop: Any = None, step: Any = 1) -> Any:
        return np.arange(start, stop, step) if stop is not None else np.arange(start)

    def linspace(self, start: Any, stop: Any, num: int = 50) -> Any:
        return np.linspace(start, stop, num)

    def transpose(self, a: Any, axes: Any = None) -> Any:
        return np.transpose(a, axes)

    def norm(self, a: Any, ord: Any = None, axis: Any = None) -> Any:
        return np.linalg.norm(a, ord=ord, axis=axis)

    def cast(self, a: Any, dtype: Any) -> Any:
        return np.asarray(a).astype(dtype)

    def choice(self, rng: Any, a: int, *, size: int, p: Any = None) -> Any:
        return rng.choice(a, size=size, p=p)

    def bincount(self, x: Any, minlength: int = 0) -> Any:
        return np.bincount(np.asarray(x, dtype=np.int64), minlength=minlength)

    def nonzero(self, x: Any) -> Any:
        return np.nonzero(x)

    def svd(self, a: Any, full_matrices: bool = False) -> Any:
        return np.linalg.svd(a, full_matrices=full_matrices)

    def eigh(self, a: Any) -> Any:
        return np.linalg.eigh(a)

    def eig(self, a: Any) -> Any:
        return np.linalg.eig(a)

    def solve(self, a: Any, b: Any, assume_a: str = "gen") -> Any:
        import scipy.linalg

        return scipy.linalg.solve(a, b, assume_a=assume_a)

    def inv(self, a: Any) -> Any:
        return np.linalg.inv(a)

    def expm(self, a: Any) -> Any:
        import scipy.linalg

        return scipy.linalg.expm(a)

    def tensordot(self, a: Any, b: Any, axes: Any = 2) -> Any:
        return np.tensordot(a, b, axes=axes)

    def rng(self, seed: Any = None) -> Any:
        return np.random.default_rng(seed)

    def normal(self, rng: Any, shape: Any, dtype: Any = None) -> Any:
        return rng.normal(size=shape).astype(dtype if dtype is not None else np.float64)

    def copy(self, a: Any) -> Any:
        return np.copy(a)

    def requires_grad(self, x: Any, flag: bool = True) -> Any:
        return x  # no autodiff on host arrays

    def detach(self, x: Any) -> Any:
        return np.asarray(x)

    def value_and_grad(self, fn: Callable, argnums: Any = 0) -> Callable:
        """(value, gradient) by central differences, step 1e-6, one pair of
        evaluations per parameter (numerics/backends/numpy_backend.py:386-446
        takes differences too)."""
        idx = (argnums,) if isinstance(argnums, int) else tuple(argnums)

        def wrapped(*args: Any, **kwargs: Any):
            val = fn(*args, **kwargs)
            grads = [_central_gradient(fn, args, kwargs, i, 1e-6) for i in idx]
            return val, (grads[0] if len(grads) == 1 else tuple(grads))

        return wrapped

    def jacfwd(self, fn: Callable, argnums: int = 0) -> Callable:
        """Jacobian by central differences (step 1e-7), shaped output + input
        (numerics/backends/numpy_backend.py:265-312)."""

        def wrapped(*args: Any, **kwargs: Any):
            x = np.asarray(args[argnums], dtype=np.float64)
            y0 = np.asarray(fn(*args, **kwargs))
            cols = []
            for i in range(x.size):
                d = np.zeros(x.size)
                d[i] = 1e-7
                lo, hi = list(args), list(args)
                lo[argnums] = (x.reshape(-1) - d).reshape(x.shape)
                hi[argnums] = (x.reshape(-1) + d).reshape(x.shape)
                cols.append((np.asarray(fn(*hi, **kwargs)) - np.asarray(fn(*lo, **kwargs))).reshape(-1)
                            / 2e-7)
            jac = np.stack(cols, axis=1) if cols else np.zeros((y0.size, 0))
            return jac.reshape(y0.shape + x.shape)

        return wrapped

    def vmap(self, fn: Callable) -> Callable:
        def mapped(*args: Any) -> Any:
            return np.stack([fn(*row) for row in zip(*args)])

        return mapped


class TorchBackend(Backend):
    """The PyTorch implementation.  array/asarray keep the autograd graph of
    tensors (a dtype change is a differentiable cast), numpy dtypes are
    mapped to torch's, and the sampling helpers return numpy arrays."""

    name = "pytorch"

    def __init__(self):
        import torch

        self.xp = torch
        self.complex64, self.complex128 = torch.complex64, torch.complex128
        self.float32, self.float64 = torch.float32, torch.float64
        self.int32, self.int64, self.int8 = torch.int32, torch.int64, torch.int8
        self.bool = torch.bool
        self.int = torch.int64

    def _dtype(self, dtype: Any) -> Any:
        torch = self.xp
        if dtype is None or isinstance(dtype, torch.dtype):
            return dtype
        named = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
                 np.dtype(np.complex64): torch.complex64, np.dtype(np.complex128): torch.complex128,
                 np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
                 np.dtype(np.int8): torch.int8, np.dtype(np.bool_): torch.bool}
        if dtype is float:
            return torch.float32
        if dtype is complex:
            return torch.complex64
        try:
            return named.get(np.dtype(dtype))
        except TypeError:
            return None

    def array(self, data: Any, dtype: Any = None) -> Any:
        return self.asarray(data, dtype)

    def asarray(self, data: Any, dtype: Any = None) -> Any:
        torch = self.xp
        td = self._dtype(dtype)
        if torch.is_tensor(data):
            return data if td is None or data.dtype == td else data.to(td)
        return torch.as_tensor(np.asarray(data) if isinstance(data, (list, tuple)) else data, dtype=td)

    def to_numpy(self, data: Any) -> np.ndarray:
        if hasattr(data, "detach"):  # conj/neg views are materialised first
            return data.detach().resolve_conj().resolve_neg().cpu().numpy()
        return np.asarray(data)

    def reshape(self, a: Any, shape: Any) -> Any:
        return self.xp.reshape(self.asarray(a), tuple(shape) if not isinstance(shape, int) else (shape,))

    def moveaxis(self, a: Any, source: int, destination: int) -> Any:
        return self.xp.movedim(a, source, destination)

    def sum(self, a: Any, axis: Any = None) -> Any:
        return self.xp.sum(a) if axis is None else self.xp.sum(a, dim=axis)

    def mean(self, a: Any, axis: Any = None) -> Any:
        return self.xp.mean(a) if axis is None else self.xp.mean(a, dim=axis)

    def zeros(self, shape: Any, dtype: Any = None) -> Any:
        return self.xp.zeros(shape, dtype=self._dtype(dtype) or self.xp.float64)

    def ones(self, shape: Any, dtype: Any = None) -> Any:
        return self.xp.ones(shape, dtype=self._dtype(dtype) or self.xp.float64)

    def eye(self, n: int, dtype: Any = None) -> Any:
        return self.xp.eye(n, dtype=self._dtype(dtype) or self.xp.float64)

    def stack(self, arrays: Any, axis: int = 0) -> Any:
        return self.xp.stack([self.asarray(x) for x in arrays], dim=axis)

    def concatenate(self, arrays: Any, axis: int = 0) -> Any:
        return self.xp.cat([self.asarray(x) for x in arrays], dim=axis)

    def arange(self, start: Any, stop: Any = None, step: Any = 1) -> Any:
        return self.xp.arange(start, stop, step) if stop is not None else self.xp.arange(start)

    def linspace(self, start: Any, stop: Any, num: int = 50) -> Any:
        return self.xp.linspace(start, stop, num)

    def transpose(self, a: Any, axes: Any = None) -> Any:
        return a.permute(*axes) if axes is not None else a.permute(*reversed(range(a.dim())))

    def norm(self, a: Any, ord: Any = None, axis: Any = None) -> Any:
        return self.xp.linalg.norm(a, ord=ord, dim=axis)

    def cast(self, a: Any, dtype: Any) -> Any:
        td = self._dtype(dtype)
        return self.asarray(a) if td is None else self.asarray(a).to(td)

    def choice(self, rng: Any, a: int, *, size: int, p: Any = None) -> Any:
        torch = self.xp
        w = torch.full((a,), 1.0 / a, dtype=torch.float64) if p is None else \
            torch.as_tensor(self.to_numpy(p), dtype=torch.float64)
        return torch.multinomial(w, size, replacement=True, generator=rng).numpy()

    def bincount(self, x: Any, minlength: int = 0) -> Any:
        return self.xp.bincount(self.xp.as_tensor(self.to_numpy(x), dtype=self.xp.int64),
                                minlength=minlength).numpy()

    def nonzero(self, x: Any) -> Any:
        return np.nonzero(self.to_numpy(x))

    def svd(self, a: Any, full_matrices: bool = False) -> Any:
        return self.xp.linalg.svd(a, full_matrices=full_matrices)

    def eigh(self, a: Any) -> Any:
        return self.xp.linalg.eigh(a)

    def eig(self, a: Any) -> Any:
        return self.xp.linalg.eig(a)

    def solve(self, a: Any, b: Any, assume_a: str = "gen") -> Any:
        return self.xp.linalg.solve(a, b)

    def inv(self, a: Any) -> Any:
        return self.xp.linalg.inv(a)

    def expm(self, a: Any) -> Any:
        return self.xp.linalg.matrix_exp(a)

    def tensordot(self, a: Any, b: Any, axes: Any = 2) -> Any:
        if isinstance(axes, int):
            return self.xp.tensordot(a, b, dims=axes)
        return self.xp.tensordot(a, b, dims=(list(axes[0]), list(axes[1])))

    def rng(self, seed: Any = None) -> Any:
        g = self.xp.Generator()
        if seed is not None:
            g.manual_seed(int(seed))
        return g

    def normal(self, rng: Any, shape: Any, dtype: Any = None) -> Any:
        return self.xp.normal(0.0, 1.0, size=tuple(shape), generator=rng,
                              dtype=self._dtype(dtype) or self.xp.float64)

    def copy(self, a: Any) -> Any:
        return self.asarray(a).clone()

    def requires_grad(self, x: Any, flag: bool = True) -> Any:
        return x.requires_grad_(flag) if hasattr(x, "requires_grad_") else x

    def detach(self, x: Any) -> Any:
        return x.detach() if hasattr(x, "detach") else x

    def value_and_grad(self, fn: Callable, argnums: Any = 0) -> Callable:
        return value_and_grad(fn, argnums)

    def jacfwd(self, fn: Callable, argnums: int = 0) -> Callable:
        from torch.func import jacfwd

        return jacfwd(fn, argnums=argnums)

    def vmap(self, fn: Callable) -> Callable:
        from torch.func import vmap

        return vmap(fn)


def _central_gradient(fn: Callable, args: Any, kwargs: dict, i: int, h: float) -> np.ndarray:
    x = np.asarray(args[i], dtype=np.float64)
    flat = x.reshape(-1)
    g = np.zeros(flat.size)
    for k in range(flat.size):
        lo, hi = list(args), list(args)
        xl, xh = flat.copy(), flat.copy()
        xl[k] -= h
        xh[k] += h
        lo[i], hi[i] = xl.reshape(x.shape), xh.reshape(x.shape)
        g[k] = (float(fn(*hi, **kwargs)) - float(fn(*lo, **kwargs))) / (2 * h)
    return g.reshape(x.shape)


def _new_backend(name: str) -> Backend:
    return NumpyBackend() if name == "numpy" else TorchBackend()


def normalize_backend_name(name: Any) -> str:
    if isinstance(name, Backend):
        return name.name
    key = str(name).lower()
    if key not in _ALIASES:
        raise ValueError(f"Unknown backend '{name}'. Supported: numpy, pytorch")
    return _ALIASES[key]


def _as_instance(name_or_instance: Any) -> Backend:
    if isinstance(name_or_instance, Backend):
        return name_or_instance
    return _new_backend(normalize_backend_name(name_or_instance))


def set_backend(name_or_instance: Any) -> Backend:
    """Select the process-wide backend by name or instance
    (numerics/context.py:18-32)."""
    inst = _as_instance(name_or_instance)
    _global["name"], _global["inst"] = inst.name, inst
    return inst


def get_backend(name: Optional[Any] = None) -> Backend:
    """The configured backend instance when ``name`` is None (the thread's
    ``use_backend`` selection first), else a new instance of the named backend,
    as numerics/api.py:179-252 returns one per call."""
    if name is not None:
        return _as_instance(name)
    inst = getattr(_state, "inst", None) or _global.get("inst")
    if inst is None:
        inst = _global["inst"] = _new_backend(_global["name"])
    return inst


class use_backend:
    """Context manager that selects a backend for the current thread
    (numerics/context.py:35-50)."""

    def __init__(self, name: Any):
        self.inst = _as_instance(name)
        self.name = self.inst.name
        self.prev: Optional[Backend] = None

    def __enter__(self) -> Backend:
        self.prev = getattr(_state, "inst", None)
        _state.inst = self.inst
        return self.inst

    def __exit__(self, *exc: Any) -> bool:
        _state.inst = self.prev
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
