"""Pauli strings and Pauli sums.

Conventions follow libs/quantum_library/kernels/pauli.py of the reference:
codes 0/1/2/3 = I/X/Y/Z (pauli.py:18-50), Y = [[0, -i], [i, 0]] (gates.py:173-177),
and Hamiltonians as lists of (coefficient, [(Pauli, qubit), ...])
(libs/circuits_library/hamiltonians.py:1-26).

On the device a Pauli string is a pair of bit masks (x, z) over amplitude
index bits, bit = n-1-qubit: X where only x is set, Z where only z is set, Y
where both are set.  The dense builders below (pauli_string_to_matrix,
pauli_string_sum_dense, heisenberg_hamiltonian) are host-side NumPy
utilities kept for API compatibility and small-n checks; the engine never
builds a 2^n x 2^n matrix.
"""
from __future__ import annotations

from typing import Any, Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .gates import PAULI

__all__ = [
    "PauliSum", "ps2xyz", "xyz2ps", "pauli_string_to_matrix", "pauli_string_sum_dense",
    "heisenberg_hamiltonian", "build_tfim_terms", "heisenberg_terms", "to_masks",
]

_CODE = {"I": 0, "X": 1, "Y": 2, "Z": 3}
_NAME = "IXYZ"


def _code_masks(ps: Sequence[int]) -> Tuple[int, int]:
    """(x, z) masks over list positions of a code list (bit i = position i):
    X sets x, Z sets z, Y sets both."""
    x = z = 0
    for i, v in enumerate(ps):
        v = int(v)
        if v in (1, 2):
            x |= 1 << i
        if v in (2, 3):
            z |= 1 << i
    return x, z


def _positions(mask: int) -> List[int]:
    out = []
    while mask:
        low = mask & -mask
        out.append(low.bit_length() - 1)
        mask ^= low
    return out


def ps2xyz(ps: Sequence[int]) -> Dict[str, List[int]]:
    """Code list (0/1/2/3 = I/X/Y/Z per position) -> {"x": [...], "y": [...],
    "z": [...]} position lists, ascending (kernels/pauli.py:18-31 contract)."""
    x, z = _code_masks(ps)
    return {"x": _positions(x & ~z), "y": _positions(x & z), "z": _positions(z & ~x)}


def xyz2ps(xyz: Dict[str, List[int]], n: Optional[int] = None) -> List[int]:
    """Inverse of ps2xyz; the length defaults to one past the largest position
    named (1 if none), unnamed positions are I (kernels/pauli.py:34-50
    contract).  A position named twice takes the later of x, y, z."""
    code = {}
    for key, val in (("x", 1), ("y", 2), ("z", 3)):
        for i in xyz.get(key, []):
            code[int(i)] = val
    length = (max(code) + 1 if code else 1) if n is None else int(n)
    return [code.get(i, 0) for i in range(length)]


def pauli_string_to_matrix(ps: Sequence[int]) -> np.ndarray:
    out = np.ones((1, 1), dtype=np.complex128)
    for code in ps:
        out = np.kron(out, PAULI[_NAME[int(code)]])
    return out


def pauli_string_sum_dense(ls: Sequence[Sequence[int]],
                           weights: Optional[Sequence[float]] = None) -> np.ndarray:
    if not ls:
        return np.zeros((1, 1), dtype=np.complex128)
    n = len(ls[0])
    H = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    weights = [1.0] * len(ls) if weights is None else weights
    for ps, w in zip(ls, weights):
        H = H + float(w) * pauli_string_to_matrix(ps)
    return H


def heisenberg_terms(num_qubits: int, edges: Sequence[Tuple[int, int]], *, hzz: float = 1.0,
                     hxx: float = 1.0, hyy: float = 1.0, hz: float = 0.0, hx: float = 0.0,
                     hy: float = 0.0) -> List[Tuple[float, List[Tuple[str, int]]]]:
    """Heisenberg model as a term list, in the order of the reference's
    heisenberg_hamiltonian (pauli.py:121-172): per edge ZZ, XX, YY; then
    per site Z, X, Y; zero couplings are skipped."""
    terms: List[Tuple[float, List[Tuple[str, int]]]] = []
    for a, b in edges:
        for c, P in ((hzz, "Z"), (hxx, "X"), (hyy, "Y")):
            if c != 0.0:
                terms.append((float(c), [(P, int(a)), (P, int(b))]))
    for q in range(num_qubits):
        for c, P in ((hz, "Z"), (hx, "X"), (hy, "Y")):
            if c != 0.0:
                terms.append((float(c), [(P, q)]))
    return terms


def heisenberg_hamiltonian(num_qubits: int, edges: Sequence[Tuple[int, int]], **kw: float):
    """Dense Heisenberg Hamiltonian (host, small n only)."""
    return PauliSum(num_qubits, heisenberg_terms(num_qubits, edges, **kw)).to_dense()


def build_tfim_terms(n: int, hzz: float = 1.0, hx: float = -1.0, pbc: bool = False):
    """hzz * sum Z_i Z_{i+1} + hx * sum X_i as a term list: the nearest-neighbour
    bonds first (with the wrap-around bond last when pbc), then the fields, the
    order of libs/circuits_library/hamiltonians.py:17-26."""
    bonds = [(i, (i + 1) % n) for i in range(n - 1 + (1 if pbc and n > 1 else 0))]
    zz = [(hzz, [("Z", a), ("Z", b)]) for a, b in bonds]
    return zz + [(hx, [("X", q)]) for q in range(n)]


def _real(c: Any) -> float:
    c = complex(c)
    if abs(c.imag) > 1e-12 * max(1.0, abs(c.real)):
        raise ValueError(f"Pauli-sum coefficient {c} is not real (operator not Hermitian)")
    return c.real


class PauliSum:
    """sum_k c_k P_k over n qubits, P_k a tensor product of single-qubit Paulis.

    ``terms`` is a list of (coeff, [(P, qubit), ...]) with P in "IXYZ"; an
    empty factor list is the identity."""

    def __init__(self, num_qubits: int, terms: Iterable[Tuple[Any, Sequence[Tuple[str, int]]]]):
        self.n = int(num_qubits)
        self.terms: List[Tuple[float, List[Tuple[str, int]]]] = []
        for c, fac in terms:
            seen = set()
            clean = []
            for P, q in fac:
                P = str(P).upper()
                q = int(q)
                if P not in _CODE:
                    raise ValueError(f"invalid Pauli symbol {P!r}")
                if not 0 <= q < self.n:
                    raise ValueError(f"qubit {q} out of range for {self.n} qubits")
                if q in seen:
                    raise ValueError(f"qubit {q} appears twice in one Pauli string")
                seen.add(q)
                if P != "I":
                    clean.append((P, q))
            self.terms.append((_real(c), clean))

    def __len__(self) -> int:
        return len(self.terms)

    @classmethod
    def from_codes(cls, ls: Sequence[Sequence[int]], weights: Optional[Sequence[float]] = None):
        n = len(ls[0]) if ls else 0
        weights = [1.0] * len(ls) if weights is None else weights
        terms = []
        for ps, w in zip(ls, weights):
            terms.append((w, [(_NAME[int(c)], q) for q, c in enumerate(ps) if int(c) != 0]))
        return cls(n, terms)

    @classmethod
    def coerce(cls, obj: Any, num_qubits: int) -> "PauliSum":
        """Accept a PauliSum, a term list [(c, [(P, q), ...])], a (codes, weights)
        pair, an OpenFermion-style {((q, P), ...): c} dict, or an object with a
        ``terms`` attribute holding one of these."""
        if isinstance(obj, PauliSum):
            if obj.n != num_qubits:
                raise ValueError(f"Pauli sum is on {obj.n} qubits, circuit has {num_qubits}")
            return obj
        if hasattr(obj, "terms") and not isinstance(obj, (list, tuple, dict)):
            return cls.coerce(obj.terms, num_qubits)
        if isinstance(obj, dict):
            terms = []
            for key, c in obj.items():
                terms.append((c, [(str(P), int(q)) for q, P in key]))
            return cls(num_qubits, terms)
        if (isinstance(obj, tuple) and len(obj) == 2 and len(obj[0]) > 0
                and isinstance(obj[0][0], (list, tuple, np.ndarray))
                and all(isinstance(v, (int, np.integer)) for v in obj[0][0])):
            ps = cls.from_codes(obj[0], obj[1])
            return cls(num_qubits, ps.terms)
        return cls(num_qubits, list(obj))

    def to_masks(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(x, z, coeff) with masks over index bits (bit = n-1-qubit)."""
        x = np.zeros(len(self.terms), dtype=np.uint64)
        z = np.zeros(len(self.terms), dtype=np.uint64)
        c = np.zeros(len(self.terms), dtype=np.float64)
        for k, (coef, fac) in enumerate(self.terms):
            xm = zm = 0
            for P, q in fac:
                b = 1 << (self.n - 1 - q)
                if P in ("X", "Y"):
                    xm |= b
                if P in ("Z", "Y"):
                    zm |= b
            x[k], z[k], c[k] = xm, zm, coef
        return x, z, c

    def to_codes(self) -> Tuple[List[List[int]], List[float]]:
        ls, ws = [], []
        for coef, fac in self.terms:
            ps = [0] * self.n
            for P, q in fac:
                ps[q] = _CODE[P]
            ls.append(ps)
            ws.append(coef)
        return ls, ws

    def to_dense(self) -> np.ndarray:
        ls, ws = self.to_codes()
        return pauli_string_sum_dense(ls, ws)


def to_masks(obj: Any, num_qubits: int):
    return PauliSum.coerce(obj, num_qubits).to_masks()
