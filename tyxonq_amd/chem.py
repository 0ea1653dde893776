"""Chemistry inputs of the UCC workload: Jordan-Wigner Pauli sums, synthetic
molecular integrals and the UCCSD ansatz as Pauli-string rotations.

The reference builds these with OpenFermion (absent here); this module
restates the textbook algebra it relies on:

* Jordan-Wigner, OpenFermion convention: a_p = (X_p + iY_p)/2 Z_{p-1}..Z_0,
  a_p^dag = (X_p - iY_p)/2 Z_{p-1}..Z_0 (libs/circuits_library/ucc.py:75-79).
* Spin-orbital index idx maps to circuit qubit n-1-idx (reverse_qop_idx,
  libs/hamiltonian_encoding/pauli_io.py:87-93); with the engine's bit order
  (qubit q on index bit n-1-q) spin orbital idx is index bit idx.
* Excitations and their shared parameters follow get_ex1_ops / get_ex2_ops
  (applications/chem/algorithms/ucc.py:680-833) in "fermion" mode: beta
  occupied 0..no-1, beta virtual no..no+nv-1, alpha occupied no+nv..,
  alpha virtual 2no+nv..
* An excitation is exp(theta (T - T^dag)); JW of T - T^dag is a sum of
  mutually commuting Pauli strings v_k P_k (v_k imaginary), applied as
  rotations exp(-i phi/2 P_k) with phi = -2 Im(v_k) theta
  (build_ucc_circuit(trotter=True), libs/circuits_library/ucc.py:84-123).
* The Hartree-Fock state sets X on qubits n-1-i (i < na) and n/2-1-i
  (i < nb) (_hf_init_ops, libs/circuits_library/ucc.py:13-26).

Pauli operators are held as {(x, z): coefficient} in the normal form
prod_q X_q^{x_q} Z_q^{z_q} over index bits; (X^a Z^b)(X^c Z^d) =
(-1)^{|b & c|} X^{a^c} Z^{b^d}, and X Z = -iY converts to Hermitian strings.
"""
from __future__ import annotations

import math
from functools import lru_cache
from typing import Dict, List, Sequence, Tuple

import numpy as np

from .circuit import Circuit

XZ = Dict[Tuple[int, int], complex]
Terms = List[Tuple[float, List[Tuple[str, int]]]]


def _popc(v: int) -> int:
    return bin(v).count("1")


def xz_mul(a: XZ, b: XZ) -> XZ:
    out: XZ = {}
    for (xa, za), ca in a.items():
        for (xb, zb), cb in b.items():
            key = (xa ^ xb, za ^ zb)
            c = ca * cb * (-1 if _popc(za & xb) & 1 else 1)
            out[key] = out.get(key, 0.0) + c
    return out


def xz_add(acc: XZ, b: XZ, scale: complex = 1.0) -> None:
    for k, c in b.items():
        acc[k] = acc.get(k, 0.0) + scale * c


def ladder(p: int, dagger: bool) -> XZ:
    """JW image of a_p (dagger=False) or a_p^dag: (X_p +- iY_p)/2 Z_{<p}."""
    e, low = 1 << p, (1 << p) - 1
    s = 0.5 if dagger else -0.5  # X - iY = X + XZ ; X + iY = X - XZ
    return {(e, low): 0.5, (e, low | e): s}


def fermion_product(word: Sequence[Tuple[int, bool]]) -> XZ:
    """JW image of a product of ladder operators [(index, is_creation), ...]."""
    out: XZ = {(0, 0): 1.0}
    for p, d in word:
        out = xz_mul(out, ladder(p, d))
    return out


def xz_to_terms(op: XZ, n: int, tol: float = 1e-12) -> Terms:
    """Hermitian Pauli strings (circuit qubits) with real coefficients; the
    imaginary parts of a Hermitian operator cancel and are checked to."""
    terms: Terms = []
    for (x, z), c in sorted(op.items()):
        c = c * (-1j) ** (_popc(x & z) % 4)  # X^x Z^z = (-i)^{|x&z|} P(x, z)
        if abs(c) <= tol:
            continue
        if abs(c.imag) > 1e-9 * max(1.0, abs(c)):
            raise ValueError("operator is not Hermitian (imaginary Pauli coefficient)")
        fac = []
        for b in range(n):
            xb, zb = (x >> b) & 1, (z >> b) & 1
            if xb or zb:
                fac.append(("Y" if xb and zb else "X" if xb else "Z", n - 1 - b))
        terms.append((float(c.real), sorted(fac, key=lambda t: t[1])))
    return terms


# ---- synthetic molecular Hamiltonian ----------------------------------------
def synthetic_integrals(n_spatial: int, seed: int = 1234):
    """h_pq symmetric with entries N(0,1)/sqrt(n); (pq|rs) with the 8-fold
    permutational symmetry, unique entries N(0, 0.1^2).  Draw order: the
    upper triangle of h row by row, then the unique (pq|rs) in the loop order
    p>=q, r>=s, pq>=rs."""
    rng = np.random.default_rng(seed)
    n = n_spatial
    h = np.zeros((n, n))
    for p in range(n):
        for q in range(p, n):
            h[p, q] = h[q, p] = rng.normal() / math.sqrt(n)
    g = np.zeros((n, n, n, n))
    pairs = [(p, q) for p in range(n) for q in range(p + 1)]
    for i, (p, q) in enumerate(pairs):
        for (r, s) in pairs[: i + 1]:
            v = rng.normal(0.0, 0.1)
            for a, b in ((p, q), (q, p)):
                for c, d in ((r, s), (s, r)):
                    g[a, b, c, d] = v
                    g[c, d, a, b] = v
    return h, g


def molecular_hamiltonian(h: np.ndarray, g: np.ndarray, e_core: float = 0.0) -> Terms:
    """H = E + sum_{pq,s} h_pq a+_{ps} a_{qs} + 1/2 sum_{pqrs,st} (pq|rs) a+_{ps} a+_{rt} a_{st} a_{qs}
    over spin orbitals p_beta = p, p_alpha = n_spatial + p, mapped by JW."""
    n_s = h.shape[0]
    n = 2 * n_s
    op: XZ = {(0, 0): complex(e_core)}

    def so(p, spin):
        return p + (n_s if spin else 0)

    for sp in (0, 1):
        for p in range(n_s):
            for q in range(n_s):
                if h[p, q] != 0.0:
                    xz_add(op, fermion_product([(so(p, sp), True), (so(q, sp), False)]), h[p, q])
    for s1 in (0, 1):
        for s2 in (0, 1):
            for p in range(n_s):
                for q in range(n_s):
                    for r in range(n_s):
                        for s in range(n_s):
                            v = g[p, q, r, s]
                            if v == 0.0:
                                continue
                            P, Q, R, S = so(p, s1), so(q, s1), so(r, s2), so(s, s2)
                            if P == R or Q == S:
                                continue  # a+ a+ or a a on one orbital vanishes
                            xz_add(op, fermion_product([(P, True), (R, True), (S, False), (Q, False)]),
                                   0.5 * v)
    return xz_to_terms(op, n)


# ---- UCCSD ansatz ------------------------------------------------------------
def uccsd_excitations(no: int, nv: int) -> Tuple[List[Tuple[int, ...]], List[int]]:
    """Excitation tuples (creation indices first) and shared parameter ids, in
    the order of get_ex1_ops then get_ex2_ops (ucc.py:680-833)."""
    ops: List[Tuple[int, ...]] = []
    ids: List[int] = []
    pid = -1
    for i in range(no):
        for a in range(nv):
            pid += 1
            ops += [(2 * no + nv + a, no + nv + i), (no + a, i)]
            ids += [pid, pid]
    nsingle = pid + 1

    def ao(i):
        return no + nv + i

    def av(a):
        return 2 * no + nv + a

    def bo(i):
        return i

    def bv(a):
        return no + a

    pid = -1
    d_ops: List[Tuple[int, ...]] = []
    d_ids: List[int] = []
    for i in range(no):
        for j in range(i):
            for a in range(nv):
                for b in range(a):
                    pid += 1
                    d_ops += [(av(b), av(a), ao(i), ao(j)), (bv(b), bv(a), bo(i), bo(j))]
                    d_ids += [pid, pid]
    for i in range(no):
        for j in range(i + 1):
            for a in range(nv):
                for b in range(a + 1):
                    if i == j and a == b:
                        pid += 1
                        d_ops.append((bv(a), av(a), ao(i), bo(i)))
                        d_ids.append(pid)
                        continue
                    pid += 1
                    d_ops += [(bv(b), av(a), ao(i), bo(j)), (av(b), bv(a), bo(i), ao(j))]
                    d_ids += [pid, pid]
                    if i != j and a != b:
                        pid += 1
                        d_ops += [(bv(a), av(b), ao(i), bo(j)), (av(a), bv(b), bo(i), ao(j))]
                        d_ids += [pid, pid]
    return ops + d_ops, ids + [nsingle + k for k in d_ids]


def excitation_strings(f_idx: Sequence[int], n: int) -> List[Tuple[complex, int, int]]:
    """JW of T - T^dag for the excitation f_idx as (v, x, z) with P(x, z) the
    Hermitian string on index bits (spin orbital idx = bit idx)."""
    return list(_excitation_strings(tuple(int(i) for i in f_idx), int(n)))


@lru_cache(maxsize=None)
def _excitation_strings(f_idx: Tuple[int, ...], n: int) -> Tuple[Tuple[complex, int, int], ...]:
    k = len(f_idx)
    if k == 2:
        word = [(f_idx[0], True), (f_idx[1], False)]
    elif k == 4:
        word = [(f_idx[0], True), (f_idx[1], True), (f_idx[2], False), (f_idx[3], False)]
    else:
        raise ValueError("excitations have 2 or 4 indices")
    t = fermion_product(word)
    td = fermion_product([(p, not d) for p, d in reversed(word)])
    op: XZ = {}
    xz_add(op, t, 1.0)
    xz_add(op, td, -1.0)
    out = []
    for (x, z), c in sorted(op.items()):
        c = c * (-1j) ** (_popc(x & z) % 4)
        if abs(c) > 1e-12:
            out.append((complex(c), x, z))
    return tuple(out)


def hf_ops(n: int, na: int, nb: int) -> List[tuple]:
    return [("x", n - 1 - i) for i in range(na)] + [("x", n // 2 - 1 - i) for i in range(nb)]


def pauli_factors(x: int, z: int, n: int) -> Tuple[Tuple[str, int], ...]:
    fac = []
    for b in range(n):
        xb, zb = (x >> b) & 1, (z >> b) & 1
        if xb or zb:
            fac.append(("Y" if xb and zb else "X" if xb else "Z", n - 1 - b))
    return tuple(sorted(fac, key=lambda t: t[1]))


def ucc_circuit(params: Sequence, n: int, n_elec: Tuple[int, int],
                ex_ops: Sequence[Tuple[int, ...]], param_ids: Sequence[int]) -> Circuit:
    """HF state, then every excitation as the rotations of its JW strings with
    angle -2 Im(v) theta_{param_id} (params may be torch tensors)."""
    c = Circuit(n, ops=hf_ops(n, n_elec[0], n_elec[1]))
    for f_idx, pid in zip(ex_ops, param_ids):
        th = params[pid]
        for v, x, z in excitation_strings(f_idx, n):
            c.ops.append(("pauli_rot", pauli_factors(x, z, n), (-2.0 * v.imag) * th))
    return c


class UCCAnsatz:
    """The UCC circuit as a tqsv_op array with the parameter structure kept:
    every rotation carries the shared parameter id of its excitation and the
    coefficient -2 Im(v) of its angle, so the adjoint sweep returns dE/dtheta
    for the shared parameters directly and new angles are a vectorised fill."""

    def __init__(self, n: int, n_elec: Tuple[int, int], ex_ops: Sequence[Tuple[int, ...]],
                 param_ids: Sequence[int]):
        from . import _native as nat
        from .gates import gate_x

        self.n = n
        self.nparams = max(param_ids) + 1 if len(param_ids) else 0
        rows = []
        xmat = np.asarray(gate_x(), dtype=np.complex128).reshape(-1)
        mflat = [0.0] * 32
        for k, z in enumerate(xmat):
            mflat[2 * k], mflat[2 * k + 1] = z.real, z.imag
        for _, q in hf_ops(n, n_elec[0], n_elec[1]):
            rows.append((nat.OP_MAT1, (n - 1 - q, -1), -1, 0, 0, 0, 0, 0.0, 0, 0, 0.0, mflat))
        self.hf_rows = len(rows)
        pid, coef = [], []
        for f_idx, p in zip(ex_ops, param_ids):
            for v, x, z in excitation_strings(f_idx, n):
                c = -2.0 * v.imag
                rows.append((nat.OP_PROT, (-1, -1), int(p), 0, 0, int(x), int(z), 0.0, 0, 0, c,
                             [0.0] * 32))
                pid.append(int(p))
                coef.append(c)
        self._ops = np.array(rows, dtype=nat.OP_DTYPE)
        self.pid = np.array(pid, dtype=np.int64)
        self.coef = np.array(coef)

    @property
    def nrot(self) -> int:
        return len(self.pid)

    def ops(self, params, out: np.ndarray | None = None) -> np.ndarray:
        """The op array for `params`.  `out`: an array returned by an earlier
        call, refilled in place (only the rotation angles change; a fresh
        2 MB copy per parameter update costs more than the update itself)."""
        th = np.asarray(params, dtype=np.float64)
        if out is None or out.shape != self._ops.shape or out.dtype != self._ops.dtype:
            out = self._ops.copy()
        out["theta"][self.hf_rows:] = self.coef * th[self.pid]
        return out
