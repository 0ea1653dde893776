"""Fused Pauli-sum expectation and H|psi> vs the CPU oracle (GPU).

Tolerances: complex64 state, FP64 reductions: per-term values to 2e-6,
energies to 1e-5 relative to the sum of |coefficients|.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import statevector as O
from tests.gpu_util import rand_state, random_pauli_terms, to_gpu, to_np

pytestmark = pytest.mark.gpu


def _ham(n, terms):
    from tyxonq_amd import _native as nat
    from tyxonq_amd.pauli import PauliSum

    x, z, c = PauliSum(n, terms).to_masks()
    return nat.Ham(n, x, z, c)


@pytest.mark.parametrize("n,maxw", [(1, 1), (2, 2), (7, 4), (11, 4), (13, 6), (16, 3), (17, 8)])
def test_expval_random_sums(gpu, n, maxw):
    rng = np.random.default_rng(n * 10 + maxw)
    v = rand_state(rng, n)
    terms = random_pauli_terms(rng, n, 60, maxw=maxw) + [(0.25, [])]
    h = _ham(n, terms)
    e, nrm, per = h.expval(to_gpu(v), per_term=True)
    e_ref, vals = O.pauli_sum_expectation(v, terms, n, per_term=True)
    np.testing.assert_allclose(per, vals, atol=2e-6)
    scale = sum(abs(c) for c, _ in terms)
    assert abs(e - e_ref) < 1e-5 * scale
    assert abs(nrm - 1.0) < 2e-6


def test_expval_config2_shape_16q(gpu):
    from tyxonq_amd import bench_workloads as W

    n = 16
    terms = W.random_2local_terms(n, 500, seed=1234)
    rng = np.random.default_rng(4)
    v = rand_state(rng, n)
    h = _ham(n, terms)
    e, nrm = h.expval(to_gpu(v))
    e_ref = O.pauli_sum_expectation(v, terms, n)
    assert abs(e - e_ref) < 1e-5
    assert abs(nrm - 1) < 2e-6


def test_expval_term_order_invariant(gpu):
    n = 14
    rng = np.random.default_rng(8)
    v = to_gpu(rand_state(rng, n))
    terms = random_pauli_terms(rng, n, 80, maxw=3)
    e1, _ = _ham(n, terms).expval(v)
    perm = rng.permutation(len(terms))
    e2, _ = _ham(n, [terms[i] for i in perm]).expval(v)
    assert abs(e1 - e2) < 1e-9


@pytest.mark.parametrize("n", [3, 12])
def test_ham_apply(gpu, n):
    import torch

    rng = np.random.default_rng(n)
    v = rand_state(rng, n)
    terms = random_pauli_terms(rng, n, 20, maxw=3)
    h = _ham(n, terms)
    out = torch.empty(1 << n, dtype=torch.complex64, device="cuda")
    h.apply(to_gpu(v), out)
    want = sum(c * O.pauli_apply(v, fac, n) for c, fac in terms)
    np.testing.assert_allclose(to_np(out), want, atol=1e-5)


def _shared_chain_terms(rng, n, groups):
    """Terms that share a flip mask and the Z string outside it (the structure
    of Jordan-Wigner strings): per group 2-6 flip qubits, a common Z chain
    outside them and several X/Y/Z letter patterns on the flip qubits."""
    terms = []
    for _ in range(groups):
        w = int(rng.integers(2, 7))
        xq = [int(q) for q in rng.choice(n, w, replace=False)]
        rest = [q for q in range(n) if q not in xq]
        zq = [int(q) for q in rng.choice(rest, int(rng.integers(0, len(rest) + 1)), replace=False)]
        for _ in range(int(rng.integers(2, 7))):
            letters = ["XY"[int(rng.integers(2))] for _ in xq]
            terms.append((float(rng.normal()), list(zip(letters, xq)) + [("Z", q) for q in zq]))
    return terms


@pytest.mark.parametrize("n", [10, 14, 21])
def test_ham_apply_shared_chains_and_new_coefficients(gpu, n):
    """H|psi> for terms that share flip masks and Z chains (the per-pattern
    coefficient tables of the apply kernel), mixed with terms that share
    nothing and diagonal terms, against the oracle; again after
    set_coeffs, at sizes up to 21 qubits."""
    import torch

    rng = np.random.default_rng(100 + n)
    v = rand_state(rng, n)
    big = n > 18
    terms = (_shared_chain_terms(rng, n, 4 if big else 12)
             + random_pauli_terms(rng, n, 4 if big else 15, maxw=3))
    terms += [(0.3, [("Z", 1)]), (-0.7, [("Z", 0), ("Z", n - 1)])]
    # enough diagonal terms for the per-index diagonal vector (rebuilt by
    # set_coeffs)
    terms += [(float(rng.normal()), [("Z", q), ("Z", (q + 3) % n)]) for q in range(10)]
    h = _ham(n, terms)
    out = torch.empty(1 << n, dtype=torch.complex64, device="cuda")
    for rnd in range(2):
        h.apply(to_gpu(v), out)
        want = sum(c * O.pauli_apply(v, fac, n) for c, fac in terms)
        np.testing.assert_allclose(to_np(out), want, atol=2e-5)
        new = rng.normal(size=len(terms))
        terms = [(float(c), fac) for c, (_, fac) in zip(new, terms)]
        h.set_coeffs([c for c, _ in terms])


def test_ham_apply_jordan_wigner_hamiltonian(gpu):
    """A Jordan-Wigner Hamiltonian: hopping terms times number operators share
    a 2-bit flip mask with Z masks one bit apart (the chunk-table blocks of
    the apply kernel), plus terms on one flip mask whose Z masks differ in
    two bits (the member loop).  Against the oracle, again after new
    coefficients."""
    import torch

    from tyxonq_amd import chem

    ns = 5
    n = 2 * ns
    hh, gg = chem.synthetic_integrals(ns, 7)
    terms = chem.molecular_hamiltonian(hh, gg)
    # one flip mask (qubits 1, 6), Z masks {2, 3} x {4, 8}: two bits apart
    for zs in ([2, 4], [3, 4], [2, 8], [3, 8], [2, 3, 4, 8]):
        terms.append((0.05 * len(zs), [("X", 1), ("Y", 6)] + [("Z", q) for q in zs]))
    rng = np.random.default_rng(3)
    v = rand_state(rng, n)
    h = _ham(n, terms)
    out = torch.empty(1 << n, dtype=torch.complex64, device="cuda")
    h.apply(to_gpu(v), out)
    want = sum(c * O.pauli_apply(v, fac, n) for c, fac in terms)
    np.testing.assert_allclose(to_np(out), want, atol=2e-5)
    c2 = rng.normal(size=len(terms))
    h.set_coeffs(c2)
    h.apply(to_gpu(v), out)
    want = sum(c * O.pauli_apply(v, fac, n) for c, (_, fac) in zip(c2, terms))
    np.testing.assert_allclose(to_np(out), want, atol=5e-5)


def test_expval_deterministic(gpu):
    n = 18
    rng = np.random.default_rng(1)
    v = to_gpu(rand_state(rng, n))
    h = _ham(n, random_pauli_terms(rng, n, 100, maxw=2))
    a = h.expval(v, per_term=True)[2]
    b = h.expval(v, per_term=True)[2]
    assert np.array_equal(a, b)
