"""ctypes binding of libtqsv (include/tqsv.h).

This module is the only place that touches the C ABI.  It loads the in-tree
``libtqsv.so`` and fails loudly if it is missing: the statevector engine has
no CPU fallback.  States are torch complex64 tensors on a HIP device; their
data pointers and the current HIP stream are handed to the library.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Dict, Optional, Sequence

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libtqsv.so"
_lib: Optional[ctypes.CDLL] = None

OP_MAT1 = 1
OP_MAT2 = 2
OP_PROT = 3
# largest state the batched adjoint kernel holds in LDS (state and lambda)
BATCH_QUBITS = 13

# numpy mirror of struct tqsv_op (include/tqsv.h)
OP_DTYPE = np.dtype(
    [
        ("kind", np.int32),
        ("t", np.int32, (2,)),
        ("param", np.int32),
        ("cmask", np.uint64),
        ("cval", np.uint64),
        ("xmask", np.uint64),
        ("zmask", np.uint64),
        ("theta", np.float64),
        ("gx", np.uint64),
        ("gz", np.uint64),
        ("gcoef", np.float64),
        ("m", np.float64, (32,)),
    ],
    align=True,
)
assert OP_DTYPE.itemsize == 336


class PlanOpts(ctypes.Structure):
    _fields_ = [
        ("tile_bits", ctypes.c_int32),
        ("min_low_bits", ctypes.c_int32),
        ("max_cost", ctypes.c_int32),
        ("fuse", ctypes.c_int32),
        ("jit", ctypes.c_int32),
        ("min_qubits", ctypes.c_int32),
    ]


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("ops_in", ctypes.c_int32),
        ("gates_after_merge", ctypes.c_int32),
        ("passes", ctypes.c_int32),
        ("phases", ctypes.c_int32),
        ("tile_bits", ctypes.c_int32),
        ("bytes_per_run", ctypes.c_int64),
        ("program_bytes", ctypes.c_int64),
        ("reverse_passes", ctypes.c_int32),
        ("reverse_single", ctypes.c_int32),
        ("generated", ctypes.c_int32),
        ("reverse_generated", ctypes.c_int32),
    ]


class HamInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("nterms", ctypes.c_int32),
        ("passes", ctypes.c_int32),
        ("phases", ctypes.c_int32),
        ("flip_groups", ctypes.c_int32),
        ("unfused_terms", ctypes.c_int32),
    ]


class NativeError(RuntimeError):
    """A libtqsv call returned an error code."""


_P = ctypes.c_void_p
_I = ctypes.c_int
_U64 = ctypes.c_uint64
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
_SIGS = {
    "tqsv_version": (_I, []),
    "tqsv_last_error": (ctypes.c_char_p, []),
    "tqsv_init_basis": (_I, [_P, _I, _U64, _P]),
    "tqsv_apply_1q": (_I, [_P, _I, _I, _DP, _P]),
    "tqsv_apply_2q": (_I, [_P, _I, _I, _I, _DP, _P]),
    "tqsv_apply_kq": (_I, [_P, _P, _I, _I, ctypes.POINTER(ctypes.c_int), _DP, _P]),
    "tqsv_apply_ctrl": (_I, [_P, _I, _U64, _U64, _I, _DP, _P]),
    "tqsv_apply_diag": (_I, [_P, _I, _I, ctypes.POINTER(ctypes.c_int), _DP, _P]),
    "tqsv_pauli_rot": (_I, [_P, _I, _U64, _U64, _D, _P]),
    "tqsv_project_z": (_I, [_P, _I, _I, _I, _DP, _P]),
    "tqsv_scale": (_I, [_P, _I, _D, _D, _P]),
    "tqsv_norm2": (_I, [_P, _I, _DP, _P]),
    "tqsv_expect_z": (_I, [_P, _I, _I, _DP, _P]),
    "tqsv_dot": (_I, [_P, _P, _I, _U64, _U64, _DP, _P]),
    "tqsv_dot_noalloc": (_I, [_P, _P, _I, _U64, _U64, _P, _P]),
    "tqsv_probs": (_I, [_P, _I, _P, _P]),
    "tqsv_plan_create": (_I, [_I, _P, _I, ctypes.POINTER(PlanOpts), ctypes.POINTER(_P)]),
    "tqsv_plan_update": (_I, [_P, _P, _I, _P]),
    "tqsv_plan_get_info": (_I, [_P, ctypes.POINTER(PlanInfo)]),
    "tqsv_plan_inspect": (_I, [_I, _P, _I, ctypes.POINTER(PlanOpts), ctypes.POINTER(PlanInfo)]),
    "tqsv_plan_emulate": (_I, [_I, _P, _I, ctypes.POINTER(PlanOpts), _P]),
    "tqsv_plan_precompile": (_I, [_I, _P, _I, ctypes.POINTER(PlanOpts), _I,
                                  ctypes.POINTER(ctypes.c_int)]),
    "tqsv_plan_run": (_I, [_P, _P, _P]),
    "tqsv_plan_destroy": (None, [_P]),
    "tqsv_apply_fused": (_I, [_P, _I, _P, _I, ctypes.POINTER(PlanOpts), _P]),
    "tqsv_ham_create": (_I, [_I, _I, _P, _P, _P, ctypes.POINTER(_P)]),
    "tqsv_ham_set_coeffs": (_I, [_P, _I, _DP, _P]),
    "tqsv_ham_precompile": (_I, [_I, _I, _P, _P, _DP, ctypes.POINTER(ctypes.c_int)]),
    "tqsv_ham_expval": (_I, [_P, _P, _DP, _DP, _DP, _P]),
    "tqsv_ham_apply": (_I, [_P, _P, _P, _P]),
    "tqsv_ham_get_info": (_I, [_P, ctypes.POINTER(HamInfo)]),
    "tqsv_ham_destroy": (None, [_P]),
    "tqsv_expval": (_I, [_P, _I, _I, _P, _P, _P, _DP, _DP, _DP, _P]),
    "tqsv_adjoint": (_I, [_P, _P, _P, _P, _I, _DP, _DP, _P]),
    "tqsv_vjp": (_I, [_P, _P, _P, _I, _DP, _P]),
    "tqsv_dm_expval": (_I, [_P, _I, _I, _P, _P, _P, _P, _DP, _P]),
    "tqsv_dm_diag_amplitudes": (_I, [_P, _I, _P, _P]),
    "tqsv_batch_create": (_I, [ctypes.POINTER(_P)]),
    "tqsv_batch_destroy": (None, [_P]),
    "tqsv_batch_adjoint": (_I, [_P, _I, _P, _P, _P, _P, _P, _DP, _P]),
    "tqsv_sampler_create": (_I, [_I, ctypes.c_int64, ctypes.POINTER(_P)]),
    "tqsv_sampler_run": (_I, [_P, _P, ctypes.c_int64, _U64, _P, _P,
                              ctypes.POINTER(ctypes.c_int64), _P]),
    "tqsv_sampler_destroy": (None, [_P]),
}


def exported_symbols() -> list[str]:
    return list(_SIGS)


def lib_path() -> Path:
    return _LIB_PATH


def load() -> ctypes.CDLL:
    """Load libtqsv.so (built in-tree by ``tyxonq_amd._build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise NativeError(
            f"{_LIB_PATH} is missing: build it with `python -m tyxonq_amd._build` "
            "(the statevector engine has no CPU fallback)"
        )
    lib = ctypes.CDLL(str(_LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = load().tqsv_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        if rc == -4:
            raise NotImplementedError(msg)
        raise NativeError(f"libtqsv error {rc}: {msg}")


# ---- torch plumbing ---------------------------------------------------------

def _torch():
    import torch

    return torch


def device() -> "object":
    torch = _torch()
    if not torch.cuda.is_available():
        raise NativeError("no HIP device is visible: the tyxonq_amd statevector engine needs a GPU")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    torch = _torch()
    return int(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> int:
    torch = _torch()
    if not isinstance(t, torch.Tensor) or t.dtype != torch.complex64 or not t.is_cuda:
        raise TypeError("expected a complex64 tensor on the GPU")
    if not t.is_contiguous():
        raise ValueError("state tensor must be contiguous")
    return int(t.data_ptr())


def _nqubits(t) -> int:
    d = int(t.numel())
    n = d.bit_length() - 1
    if d < 2 or (1 << n) != d:
        raise ValueError(f"state length {d} is not a power of two >= 2")
    return n


def _dbuf(values: Sequence[float]):
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    return arr


def new_state(n: int, index: int = 0):
    torch = _torch()
    psi = torch.empty(1 << n, dtype=torch.complex64, device=device())
    init_basis(psi, index)
    return psi


def init_basis(psi, index: int = 0) -> None:
    _check(load().tqsv_init_basis(_ptr(psi), _nqubits(psi), int(index), stream_ptr()))


def mat_buffer(m) -> "ctypes.Array":
    m = np.asarray(m, dtype=np.complex128).reshape(-1)
    flat = np.empty(2 * m.size)
    flat[0::2] = m.real
    flat[1::2] = m.imag
    return _dbuf(flat)


def apply_1q(psi, bit: int, U) -> None:
    _check(load().tqsv_apply_1q(_ptr(psi), _nqubits(psi), int(bit), mat_buffer(U), stream_ptr()))


def apply_2q(psi, bit_hi: int, bit_lo: int, U) -> None:
    _check(
        load().tqsv_apply_2q(
            _ptr(psi), _nqubits(psi), int(bit_hi), int(bit_lo), mat_buffer(U), stream_ptr()
        )
    )


def apply_kq(psi, bits: Sequence[int], U):
    """U psi for a 2^k x 2^k matrix on index bits `bits` (bits[0] the most
    significant index of U); returns a new state."""
    import torch

    k = len(bits)
    m = np.ascontiguousarray(np.asarray(U, dtype=np.complex128).reshape(1 << k, 1 << k))
    buf = np.ascontiguousarray(m.view(np.float64).reshape(-1))
    b = (ctypes.c_int * k)(*[int(x) for x in bits])
    out = torch.empty_like(psi)
    _check(load().tqsv_apply_kq(_ptr(psi), _ptr(out), _nqubits(psi), k, b,
                                buf.ctypes.data_as(_DP), stream_ptr()))
    return out


def apply_ctrl(psi, ctrl_mask: int, ctrl_val: int, bit: int, U) -> None:
    _check(
        load().tqsv_apply_ctrl(
            _ptr(psi), _nqubits(psi), int(ctrl_mask), int(ctrl_val), int(bit), mat_buffer(U),
            stream_ptr(),
        )
    )


def apply_diag(psi, bits: Sequence[int], phases) -> None:
    bits = [int(b) for b in bits]
    arr = (ctypes.c_int * len(bits))(*bits)
    _check(
        load().tqsv_apply_diag(
            _ptr(psi), _nqubits(psi), len(bits), arr, mat_buffer(phases), stream_ptr()
        )
    )


def pauli_rot(psi, xmask: int, zmask: int, theta: float) -> None:
    _check(
        load().tqsv_pauli_rot(
            _ptr(psi), _nqubits(psi), int(xmask), int(zmask), float(theta), stream_ptr()
        )
    )


def project_z(psi, bit: int, keep: int) -> float:
    out = ctypes.c_double(0.0)
    _check(
        load().tqsv_project_z(
            _ptr(psi), _nqubits(psi), int(bit), int(keep), ctypes.byref(out), stream_ptr()
        )
    )
    return out.value


def scale(psi, value: complex) -> None:
    v = complex(value)
    _check(load().tqsv_scale(_ptr(psi), _nqubits(psi), v.real, v.imag, stream_ptr()))


def norm2(psi) -> float:
    out = ctypes.c_double(0.0)
    _check(load().tqsv_norm2(_ptr(psi), _nqubits(psi), ctypes.byref(out), stream_ptr()))
    return out.value


def expect_z(psi, bit: int) -> float:
    out = ctypes.c_double(0.0)
    _check(
        load().tqsv_expect_z(_ptr(psi), _nqubits(psi), int(bit), ctypes.byref(out), stream_ptr())
    )
    return out.value


def dot(u, v, xmask: int = 0, zmask: int = 0) -> complex:
    """sum_i (-1)^{|i&z|} conj(u[i]) v[i^x], accumulated in FP64."""
    out = (ctypes.c_double * 2)()
    n = _nqubits(u)
    if _nqubits(v) != n:
        raise ValueError("state sizes differ")
    _check(load().tqsv_dot(_ptr(u), _ptr(v), n, int(xmask), int(zmask), out, stream_ptr()))
    return complex(out[0], out[1])


def dot_noalloc(u, v, out_dev, xmask: int = 0, zmask: int = 0) -> None:
    """dot() into out_dev[0:2] (a float64 device tensor), allocating nothing
    and without synchronising (tqsv_dot_noalloc)."""
    n = _nqubits(u)
    if _nqubits(v) != n:
        raise ValueError("state sizes differ")
    if (out_dev.numel() < 2 or not out_dev.is_cuda or str(out_dev.dtype) != "torch.float64"
            or not out_dev.is_contiguous()):
        raise ValueError("out_dev must be a contiguous float64 device tensor of >= 2 elements")
    _check(load().tqsv_dot_noalloc(_ptr(u), _ptr(v), n, int(xmask), int(zmask),
                                   int(out_dev.data_ptr()), stream_ptr()))


def probs(psi):
    torch = _torch()
    out = torch.empty(psi.numel(), dtype=torch.float64, device=psi.device)
    _check(load().tqsv_probs(_ptr(psi), _nqubits(psi), int(out.data_ptr()), stream_ptr()))
    return out


# ---- plans -----------------------------------------------------------------

def _opts(tile_bits: int = 0, min_low_bits: int = 0, max_cost: int = 0, fuse: bool = True,
          jit: int = 0, min_qubits: int = 0):
    return PlanOpts(int(tile_bits), int(min_low_bits), int(max_cost), 1 if fuse else 0, int(jit),
                    int(min_qubits))


def inspect(n: int, ops: np.ndarray, *, tile_bits: int = 0, min_low_bits: int = 0,
            max_cost: int = 0, fuse: bool = True) -> dict:
    """Plan on the host only and return the plan statistics (no GPU needed)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    o = _opts(tile_bits, min_low_bits, max_cost, fuse)
    inf = PlanInfo()
    _check(load().tqsv_plan_inspect(int(n), ops.ctypes.data if len(ops) else None, len(ops),
                                    ctypes.byref(o), ctypes.byref(inf)))
    return {f: getattr(inf, f) for f, _ in PlanInfo._fields_}


def precompile(n: int, ops: np.ndarray, *, adjoint: bool = False, tile_bits: int = 0,
               min_low_bits: int = 0, max_cost: int = 0, fuse: bool = True,
               min_qubits: int = 0) -> int:
    """Plan on the host and compile the generated kernels of the plan (and,
    with adjoint=True, of its reverse sweep) into the compiler's disk cache,
    without a GPU, so that a later Plan(..., jit=1) of the same structure
    starts without compiling.  Returns the number of distinct kernels."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    o = _opts(tile_bits, min_low_bits, max_cost, fuse, 1, min_qubits)
    k = ctypes.c_int(0)
    _check(load().tqsv_plan_precompile(int(n), ops.ctypes.data if len(ops) else None, len(ops),
                                       ctypes.byref(o), 1 if adjoint else 0, ctypes.byref(k)))
    return int(k.value)


def ham_precompile(n: int, xmask, zmask, coeff) -> int:
    """Plan a Pauli sum's expectation passes on the host and compile the
    generated kernels Ham(n, ...) would use into the compiler's disk cache,
    without a GPU.  Returns the number of distinct kernels."""
    x = np.ascontiguousarray(xmask, dtype=np.uint64)
    z = np.ascontiguousarray(zmask, dtype=np.uint64)
    c = np.ascontiguousarray(coeff, dtype=np.float64)
    k = ctypes.c_int(0)
    _check(load().tqsv_ham_precompile(int(n), len(c), x.ctypes.data if len(c) else None,
                                      z.ctypes.data if len(c) else None,
                                      c.ctypes.data_as(_DP) if len(c) else None, ctypes.byref(k)))
    return int(k.value)


def emulate(n: int, ops: np.ndarray, state: np.ndarray, *, tile_bits: int = 0,
            min_low_bits: int = 0, max_cost: int = 0, fuse: bool = True) -> np.ndarray:
    """Run the planned passes on a host complex64 copy of `state` (no GPU)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    st = np.ascontiguousarray(np.asarray(state, dtype=np.complex64).copy())
    o = _opts(tile_bits, min_low_bits, max_cost, fuse)
    _check(load().tqsv_plan_emulate(int(n), ops.ctypes.data if len(ops) else None, len(ops),
                                    ctypes.byref(o), st.ctypes.data))
    return st


class Plan:
    """A planned (fused) circuit: owns its device program."""

    def __init__(self, n: int, ops: np.ndarray, *, tile_bits: int = 0, min_low_bits: int = 0,
                 max_cost: int = 0, fuse: bool = True, jit: int = 0, min_qubits: int = 0):
        load()
        self.n = int(n)
        self._ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        self._h = ctypes.c_void_p()
        o = _opts(tile_bits, min_low_bits, max_cost, fuse, jit, min_qubits)
        _check(
            load().tqsv_plan_create(
                self.n, self._ops.ctypes.data if len(self._ops) else None, len(self._ops),
                ctypes.byref(o), ctypes.byref(self._h),
            )
        )
        self.nparams = int(self._ops["param"].max()) + 1 if len(self._ops) else 0

    def update(self, ops: np.ndarray) -> None:
        self._ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        _check(
            load().tqsv_plan_update(
                self._h, self._ops.ctypes.data if len(self._ops) else None, len(self._ops),
                stream_ptr(),
            )
        )

    @property
    def ops(self) -> np.ndarray:
        """The op rows the plan currently holds (replaced, never mutated, by update)."""
        return self._ops

    def info(self) -> dict:
        inf = PlanInfo()
        _check(load().tqsv_plan_get_info(self._h, ctypes.byref(inf)))
        return {f: getattr(inf, f) for f, _ in PlanInfo._fields_}

    def run(self, psi) -> None:
        if _nqubits(psi) != self.n:
            raise ValueError(f"plan is for {self.n} qubits, state has {_nqubits(psi)}")
        _check(load().tqsv_plan_run(self._h, _ptr(psi), stream_ptr()))

    def adjoint(self, ham: "Ham", psi, lam, nparams: Optional[int] = None):
        """Forward-run psi, then the adjoint sweep.  Returns (energy, grad[nparams])."""
        npar = self.nparams if nparams is None else int(nparams)
        grad = np.zeros(max(npar, 1))
        e = ctypes.c_double(0.0)
        _check(
            load().tqsv_adjoint(
                self._h, ham._h, _ptr(psi), _ptr(lam), npar,
                grad.ctypes.data_as(_DP), ctypes.byref(e), stream_ptr(),
            )
        )
        return e.value, grad[:npar]

    def vjp(self, psi, lam, nparams: Optional[int] = None) -> np.ndarray:
        npar = self.nparams if nparams is None else int(nparams)
        grad = np.zeros(max(npar, 1))
        _check(
            load().tqsv_vjp(self._h, _ptr(psi), _ptr(lam), npar, grad.ctypes.data_as(_DP),
                            stream_ptr())
        )
        return grad[:npar]

    def close(self) -> None:
        if self._h:
            load().tqsv_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ham:
    """A Pauli sum planned for n-qubit expectations (owns device buffers)."""

    def __init__(self, n: int, xmask: Sequence[int], zmask: Sequence[int], coeff: Sequence[float]):
        load()
        self.n = int(n)
        self.x = np.ascontiguousarray(np.asarray(xmask, dtype=np.uint64))
        self.z = np.ascontiguousarray(np.asarray(zmask, dtype=np.uint64))
        self.c = np.ascontiguousarray(np.asarray(coeff, dtype=np.float64))
        if not (len(self.x) == len(self.z) == len(self.c)):
            raise ValueError("xmask, zmask and coeff must have equal length")
        self._h = ctypes.c_void_p()
        nt = len(self.c)
        _check(
            load().tqsv_ham_create(
                self.n, nt, self.x.ctypes.data if nt else None, self.z.ctypes.data if nt else None,
                self.c.ctypes.data if nt else None, ctypes.byref(self._h),
            )
        )

    @property
    def nterms(self) -> int:
        return len(self.c)

    def set_coeffs(self, coeff: Sequence[float]) -> None:
        """New coefficients for the same terms (plans and device data kept)."""
        c = np.ascontiguousarray(np.asarray(coeff, dtype=np.float64))
        if c.shape != (self.nterms,):
            raise ValueError(f"expected {self.nterms} coefficients, got {c.shape}")
        _check(load().tqsv_ham_set_coeffs(self._h, self.nterms, c.ctypes.data_as(_DP), stream_ptr()))
        self.c = c

    def expval(self, psi, per_term: bool = False):
        """Returns (energy, norm2) or (energy, norm2, terms)."""
        if _nqubits(psi) != self.n:
            raise ValueError(f"Hamiltonian is for {self.n} qubits, state has {_nqubits(psi)}")
        terms = np.zeros(max(self.nterms, 1))
        e = ctypes.c_double(0.0)
        nrm = ctypes.c_double(0.0)
        _check(
            load().tqsv_ham_expval(
                self._h, _ptr(psi), terms.ctypes.data_as(_DP), ctypes.byref(e),
                ctypes.byref(nrm), stream_ptr(),
            )
        )
        if per_term:
            return e.value, nrm.value, terms[: self.nterms]
        return e.value, nrm.value

    def apply(self, psi, out) -> None:
        _check(load().tqsv_ham_apply(self._h, _ptr(psi), _ptr(out), stream_ptr()))

    def info(self) -> dict:
        inf = HamInfo()
        _check(load().tqsv_ham_get_info(self._h, ctypes.byref(inf)))
        return {f: getattr(inf, f) for f, _ in HamInfo._fields_}

    def close(self) -> None:
        if self._h:
            load().tqsv_ham_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def version() -> int:
    return int(load().tqsv_version())


if os.environ.get("TQSV_EAGER_LOAD"):
    load()


class Sampler:
    """Shot sampler for n-qubit states (owns its device scratch)."""

    def __init__(self, n: int, max_shots: int):
        load()
        self.n = int(n)
        self.max_shots = int(max_shots)
        self._h = ctypes.c_void_p()
        _check(load().tqsv_sampler_create(self.n, self.max_shots, ctypes.byref(self._h)))
        self._pin = None  # pinned (index, count) staging, allocated on first use

    def _staging(self, cap: int):
        """Page-locked host buffers for the (index, count) results, so that the
        device-to-host copies run as direct DMA (allocated once per sampler,
        sized for its largest result)."""
        if self._pin is None or self._pin[0].shape[0] < cap:
            import torch

            size = max(cap, min(self.max_shots, 1 << self.n))
            ti = torch.empty(size, dtype=torch.int64, pin_memory=True)
            tc = torch.empty(size, dtype=torch.int64, pin_memory=True)
            self._pin = (ti.numpy().view(np.uint64), tc.numpy(), ti, tc)
        return self._pin[0], self._pin[1]

    def sample(self, psi, shots: int, seed: int = 0):
        """(index, count) arrays of the distinct sampled basis states, index
        ascending (uint64 / int64)."""
        if _nqubits(psi) != self.n:
            raise ValueError(f"sampler is for {self.n} qubits, state has {_nqubits(psi)}")
        shots = int(shots)
        cap = max(1, min(shots, 1 << self.n))
        idx, cnt = self._staging(cap)
        nu = ctypes.c_int64(0)
        _check(load().tqsv_sampler_run(self._h, _ptr(psi), shots, int(seed) & ((1 << 64) - 1),
                                       idx.ctypes.data, cnt.ctypes.data, ctypes.byref(nu),
                                       stream_ptr()))
        k = int(nu.value)
        # the staging is reused by the next call: the results are copied out
        return idx[:k].copy(), cnt[:k].copy()

    def close(self) -> None:
        if self._h:
            load().tqsv_sampler_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_samplers: Dict[int, "Sampler"] = {}


def sample_counts(psi, shots: int, seed: int = 0):
    """Sample `shots` basis states of psi on the GPU (a cached sampler per
    qubit count, grown to the largest shot count seen)."""
    n = _nqubits(psi)
    s = _samplers.get(n)
    if s is None or s.max_shots < int(shots):
        if s is not None:
            s.close()
        s = Sampler(n, max(int(shots), 1024))
        _samplers[n] = s
    return s.sample(psi, shots, seed)


class Batch:
    """Energy and gradient of several small requests (at most 13 qubits, plans
    built with min_qubits=13) in one kernel launch."""

    def __init__(self):
        load()
        self._h = ctypes.c_void_p()
        _check(load().tqsv_batch_create(ctypes.byref(self._h)))

    def adjoint(self, plans: Sequence["Plan"], hams: Sequence["Ham"], psis, nparams: Sequence[int]):
        """[(energy, grad)] per request; psis are the initial states (device)."""
        k = len(plans)
        if not (len(hams) == len(psis) == len(nparams) == k):
            raise ValueError("plans, hams, psis and nparams must have equal length")
        if k == 0:
            return []
        P = (ctypes.c_void_p * k)(*[p._h.value for p in plans])
        H = (ctypes.c_void_p * k)(*[h._h.value for h in hams])
        S = (ctypes.c_void_p * k)(*[_ptr(x) for x in psis])
        npar = [int(x) for x in nparams]
        N = (ctypes.c_int * k)(*npar)
        grads = [np.zeros(max(m, 1)) for m in npar]
        G = (ctypes.c_void_p * k)(*[g.ctypes.data for g in grads])
        E = np.zeros(k)
        _check(load().tqsv_batch_adjoint(self._h, k, ctypes.cast(P, _P), ctypes.cast(H, _P),
                                         ctypes.cast(S, _P), ctypes.cast(N, _P),
                                         ctypes.cast(G, _P), E.ctypes.data_as(_DP), stream_ptr()))
        return [(float(E[i]), grads[i][:npar[i]]) for i in range(k)]

    def close(self) -> None:
        if self._h:
            load().tqsv_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dm_expval(vec, n: int, xmask, zmask, coeff=None):
    """(sum_t coeff_t Tr(rho P_t), [Tr(rho P_t)]) of a vectorised n-qubit
    density matrix (a 2n-qubit complex64 state on the GPU)."""
    if _nqubits(vec) != 2 * int(n):
        raise ValueError("vec must hold a density matrix of n qubits (2n-qubit state)")
    x = np.ascontiguousarray(np.asarray(xmask, dtype=np.uint64))
    z = np.ascontiguousarray(np.asarray(zmask, dtype=np.uint64))
    c = None if coeff is None else np.ascontiguousarray(np.asarray(coeff, dtype=np.float64))
    terms = np.zeros(max(len(x), 1))
    val = ctypes.c_double(0.0)
    _check(load().tqsv_dm_expval(_ptr(vec), int(n), len(x), x.ctypes.data if len(x) else None,
                                 z.ctypes.data if len(z) else None,
                                 None if c is None else c.ctypes.data, terms.ctypes.data,
                                 ctypes.byref(val), stream_ptr()))
    return val.value, terms[:len(x)]


def dm_diag_amplitudes(vec, n: int):
    """n-qubit complex64 state with |amplitude|^2 = diag(rho)."""
    torch = _torch()
    out = torch.empty(1 << int(n), dtype=torch.complex64, device=vec.device)
    _check(load().tqsv_dm_diag_amplitudes(_ptr(vec), int(n), _ptr(out), stream_ptr()))
    return out

