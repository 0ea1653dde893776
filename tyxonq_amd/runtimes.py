"""Chemistry numeric runtimes on the fused expectation and adjoint path.

Mirrors the statevector engines of
applications/chem/runtimes/hea_numeric_runtime.py:15-110 (HEANumericRuntime)
and applications/chem/runtimes/ucc_numeric_runtime.py:40-330
(UCCNumericRuntime): the same constructors and ``energy`` /
``energy_and_grad`` methods.  The reference builds a dense 2^n x 2^n
Hamiltonian (hea_numeric_runtime.py:38-57, pauli_string_sum_dense) or applies
the qubit operator to a dense vector (ucc_numeric_runtime.py:286-292,
apply_op) and differentiates through autograd or finite differences.  Here
the Hamiltonian is a Pauli sum planned once for the fused FP64 expectation,
and the gradient is one adjoint sweep on the GPU.

Index conventions.  HEA Hamiltonians are term lists on circuit qubits, as in
the reference.  A UCC ``h_qubit_op`` in OpenFermion form ({((p, "X"), ...):
c} or an object with such a ``terms`` dict) indexes spin-orbitals: orbital p
is occupation bit p of the basis index, i.e. circuit qubit n-1-p, the
Jordan-Wigner layout of tyxonq_amd.chem.
"""
from __future__ import annotations

from typing import Any, Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from . import chem
from .pauli import PauliSum

Terms = List[Tuple[float, List[Tuple[str, int]]]]


def _engine_name(name: Optional[str]) -> str:
    name = (name or "statevector").lower()
    if name != "statevector":
        raise NotImplementedError(f"numeric engine {name!r} is not available (statevector only)")
    return name


class HEANumericRuntime:
    """Energy and gradient of a hardware-efficient ansatz built by
    ``get_circuit(params)`` (hea_numeric_runtime.py:15-110)."""

    def __init__(self, n: int, layers: int, hamiltonian: Terms, numeric_engine: Optional[str] = None,
                 *, circuit_template: Optional[list] = None, qop: Any = None):
        if circuit_template is not None:
            raise NotImplementedError("circuit templates are not supported; pass get_circuit")
        self.n = int(n)
        self.layers = int(layers)
        self.hamiltonian = list(hamiltonian)
        self.numeric_engine = _engine_name(numeric_engine)
        self._qop_cached = qop
        from .engine import StatevectorEngine

        self._engine = StatevectorEngine("pytorch", reference_dispatch=False)

    def energy(self, params: Sequence[float], get_circuit: Callable[[Any], Any]):
        """<psi(params)|H|psi(params)>; a torch scalar that is differentiable
        when params are tensors requiring grad (hea_numeric_runtime.py:74-89)."""
        return self._engine.energy(get_circuit(params), self.hamiltonian)

    def energy_and_grad(self, params: Sequence[float],
                        get_circuit: Callable[[Any], Any]) -> Tuple[float, np.ndarray]:
        """(E, dE/dparams) by one forward and one adjoint sweep, with torch
        autograd mapping the gate angles back to ``params``
        (hea_numeric_runtime.py:91-110)."""
        import torch

        th = torch.tensor(np.asarray(params, dtype=np.float64), dtype=torch.float64,
                          requires_grad=True)
        with torch.enable_grad():
            e = self.energy(th, get_circuit)
            (g,) = torch.autograd.grad(e, th, allow_unused=True)
        grad = np.zeros(th.numel()) if g is None else g.detach().cpu().numpy().astype(np.float64)
        return float(e.detach()), grad


def spin_orbital_terms(h_qubit_op: Any, n_qubits: int) -> Terms:
    """OpenFermion-style operator on spin-orbitals -> term list on circuit
    qubits (orbital p -> qubit n-1-p)."""
    obj = h_qubit_op.terms if hasattr(h_qubit_op, "terms") and not isinstance(
        h_qubit_op, (list, tuple, dict)) else h_qubit_op
    if isinstance(obj, dict):
        out: Terms = []
        for key, c in obj.items():
            cc = complex(c)
            if abs(cc.imag) > 1e-12:
                raise ValueError("the qubit operator must be Hermitian with real coefficients")
            out.append((cc.real, [(str(P).upper(), n_qubits - 1 - int(p)) for p, P in key]))
        return out
    return [(float(c), [(str(P).upper(), n_qubits - 1 - int(p)) for P, p in fac]) for c, fac in obj]


class UCCNumericRuntime:
    """Energy and gradient of a UCC ansatz on the statevector engine
    (ucc_numeric_runtime.py:40-330): Hartree-Fock reference, every excitation
    exp(theta (T - T^dag)) as the Pauli rotations of its Jordan-Wigner
    strings, parameters shared through ``param_ids``."""

    def __init__(self, n_qubits: int, n_elec_s: Tuple[int, int], h_qubit_op: Any, *,
                 ex_ops: Optional[List[Tuple]] = None, param_ids: Optional[List[int]] = None,
                 init_state: Any = None, mode: str = "fermion", trotter: bool = False,
                 decompose_multicontrol: bool = False, numeric_engine: Optional[str] = None,
                 hamiltonian: Any = None):
        self.n_qubits = int(n_qubits)
        self.n_elec_s = (int(n_elec_s[0]), int(n_elec_s[1]))
        self.h_qubit_op = h_qubit_op
        self.ex_ops = list(ex_ops) if ex_ops is not None else None
        self.param_ids = list(param_ids) if param_ids is not None else None
        if self.ex_ops is not None and not self.param_ids:
            self.param_ids = list(range(len(self.ex_ops)))
        if init_state is not None:
            raise NotImplementedError("only the Hartree-Fock reference state is supported")
        if str(mode) != "fermion":
            raise NotImplementedError(f"mode {mode!r} is not supported (fermion only)")
        self.mode = "fermion"
        self.trotter = bool(trotter)
        self.decompose_multicontrol = bool(decompose_multicontrol)
        self.numeric_engine = _engine_name(numeric_engine)
        self.hamiltonian = hamiltonian
        self.n_params = (max(self.param_ids) + 1) if self.ex_ops else 0
        terms = spin_orbital_terms(h_qubit_op, self.n_qubits)
        x, z, c = PauliSum(self.n_qubits, terms).to_masks()
        self._ham = nat.Ham(self.n_qubits, x, z, c)
        self._ansatz = chem.UCCAnsatz(self.n_qubits, self.n_elec_s, self.ex_ops or [],
                                      self.param_ids or [])
        self._plan: Optional[nat.Plan] = None
        self._opbuf: Optional[np.ndarray] = None  # op array refilled per parameter update

    def _params(self, params) -> np.ndarray:
        if params is None:
            return np.zeros(self.n_params)
        if hasattr(params, "detach"):
            params = params.detach().cpu().numpy()
        p = np.asarray(params, dtype=np.float64).reshape(-1)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        return p

    def _run(self, params):
        import torch

        ops = self._opbuf = self._ansatz.ops(self._params(params), out=self._opbuf)
        if self._plan is None:
            self._plan = nat.Plan(self.n_qubits, ops, max_cost=100000)
        else:
            self._plan.update(ops)
        psi = nat.new_state(self.n_qubits, 0)
        lam = torch.empty_like(psi)
        return self._plan.adjoint(self._ham, psi, lam, max(self.n_params, 1))

    def energy(self, params: Optional[Sequence[float]] = None) -> float:
        """<psi(params)|H|psi(params)> (ucc_numeric_runtime.py:286-292)."""
        return float(self._run(params)[0])

    def energy_and_grad(self, params: Optional[Sequence[float]] = None, *,
                        gradient_method: str = "analytic") -> Tuple[float, np.ndarray]:
        """(E, dE/dparams) (ucc_numeric_runtime.py:294-330); "analytic" is the
        adjoint sweep, "fd" central differences of the same energies.  The
        reference differences with a step of 1e-5 on float64 states; the
        energies here carry the complex64 rounding of the state (about 1e-7),
        so the step is 1e-3 (difference noise about 1e-4, truncation about
        1e-6)."""
        if gradient_method == "fd":
            base = self._params(params)
            e0 = self.energy(base)
            g = np.zeros_like(base)
            eps = 1e-3
            for i in range(base.size):
                d = np.zeros_like(base)
                d[i] = eps
                g[i] = (self.energy(base + d) - self.energy(base - d)) / (2 * eps)
            return e0, g
        e, g = self._run(params)
        return float(e), np.asarray(g[:self.n_params], dtype=np.float64)
