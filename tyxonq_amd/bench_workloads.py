"""Synthetic workloads of the statevector benchmarks (seeded, reproducible).

Random numbers come from numpy.random.default_rng(seed), drawn in the order
each docstring lists.  The circuits use the reference's op-tuple IR, so the
same objects run on the GPU engine and on the CPU oracle.
"""
from __future__ import annotations

import math
from typing import Any, List, Sequence, Tuple

import numpy as np

from .circuit import Circuit

Terms = List[Tuple[float, List[Tuple[str, int]]]]


def brickwork_circuit(n: int = 30, layers: int = 20, seed: int = 1234) -> Circuit:
    """Per layer l: angles a[0..n) then b[0..n) ~ U[0, 2pi); RX(a_q) then RZ(b_q)
    on every qubit q; then CZ(2i, 2i+1) on even l, CZ(2i+1, 2i+2) on odd l."""
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for l in range(layers):
        a = rng.uniform(0.0, 2.0 * math.pi, size=n)
        b = rng.uniform(0.0, 2.0 * math.pi, size=n)
        for q in range(n):
            c.rx(q, float(a[q]))
            c.rz(q, float(b[q]))
        start = 0 if l % 2 == 0 else 1
        for i in range(start, n - 1, 2):
            c.cz(i, i + 1)
    return c


def random_2local_terms(n: int = 30, nterms: int = 500, seed: int = 1234) -> Terms:
    """Each term: a qubit pair (i < j) uniform over the n(n-1)/2 pairs, then a
    Pauli pair uniform over {X,Y,Z}^2, then a coefficient ~ N(0, 1);
    duplicates are kept.  The generator is seeded with seed + 1 so that the
    circuit and the Hamiltonian draws are independent."""
    rng = np.random.default_rng(seed + 1)
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    terms: Terms = []
    for _ in range(nterms):
        i, j = pairs[int(rng.integers(len(pairs)))]
        pp = int(rng.integers(9))
        coef = float(rng.normal())
        terms.append((coef, [("XYZ"[pp // 3], i), ("XYZ"[pp % 3], j)]))
    return terms


def tfim_terms(n: int, g: float = 1.0, periodic: bool = False) -> Terms:
    """H = -sum Z_i Z_{i+1} - g sum X_i (open chain unless periodic)."""
    terms: Terms = [(-1.0, [("Z", i), ("Z", i + 1)]) for i in range(n - 1)]
    if periodic and n > 2:
        terms.append((-1.0, [("Z", n - 1), ("Z", 0)]))
    terms += [(-float(g), [("X", i)]) for i in range(n)]
    return terms


def heisenberg_ring_terms(n: int, hz: float = 0.5) -> Terms:
    """H = sum_i (X_i X_{i+1} + Y_i Y_{i+1} + Z_i Z_{i+1}) + hz sum Z_i, periodic."""
    terms: Terms = []
    for i in range(n):
        j = (i + 1) % n
        for P in "XYZ":
            terms.append((1.0, [(P, i), (P, j)]))
    terms += [(float(hz), [("Z", i)]) for i in range(n)]
    return terms


def hea_ry_cx(n: int = 12, layers: int = 4, params: Any = None, seed: int = 1234) -> Tuple[Circuit, Any]:
    """layers x [RY(theta) on every qubit, CX ladder 0->1->...->n-1];
    theta ~ U[0, 2pi) drawn from the seed when params is None."""
    if params is None:
        params = np.random.default_rng(seed).uniform(0.0, 2.0 * math.pi, size=n * layers)
    c = Circuit(n)
    k = 0
    for _ in range(layers):
        for q in range(n):
            c.ry(q, params[k])
            k += 1
        for q in range(n - 1):
            c.cx(q, q + 1)
    return c, params


def ryrz_ring_ansatz(n: int = 22, layers: int = 12, params: Any = None,
                     seed: int = 1234) -> Tuple[Circuit, Any]:
    """layers x [RY on every qubit, RZ on every qubit, CX(q, q+1 mod n)];
    theta ~ N(0, 0.5^2) drawn from the seed when params is None."""
    if params is None:
        params = np.random.default_rng(seed).normal(0.0, 0.5, size=2 * n * layers)
    c = Circuit(n)
    k = 0
    for _ in range(layers):
        for q in range(n):
            c.ry(q, params[k])
            k += 1
        for q in range(n):
            c.rz(q, params[k])
            k += 1
        for q in range(n):
            c.cx(q, (q + 1) % n)
    return c, params
