"""Chemistry runtime mirrors on the GPU against the complex128 oracle."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import statevector as O
from tyxonq_amd import chem
from tyxonq_amd import runtimes as R

pytestmark = pytest.mark.gpu


def _tfim(n, g=0.9):
    t = [(-1.0, [("Z", i), ("Z", i + 1)]) for i in range(n - 1)]
    return t + [(-g, [("X", i)]) for i in range(n)]


def test_hea_runtime_energy_and_grad(gpu):
    import tyxonq_amd as tq

    n, layers = 6, 3
    H = _tfim(n)

    def get_circuit(p):
        c = tq.Circuit(n)
        k = 0
        for _ in range(layers):
            for q in range(n):
                c.ry(q, p[k])
                k += 1
            for q in range(n - 1):
                c.cx(q, q + 1)
        return c

    def oracle_e(p):
        ops, k = [], 0
        for _ in range(layers):
            for q in range(n):
                ops.append(("ry", q, float(p[k])))
                k += 1
            for q in range(n - 1):
                ops.append(("cx", q, q + 1))
        psi, _ = O.simulate(n, ops)
        return O.pauli_sum_expectation(psi, H, n)

    rt = R.HEANumericRuntime(n, layers, H)
    p = np.random.default_rng(1).uniform(0, 2 * np.pi, n * layers)
    e, g = rt.energy_and_grad(p, get_circuit)
    assert abs(e - oracle_e(p)) < 1e-6
    for j in (0, 7, n * layers - 1):
        d = np.zeros_like(p)
        d[j] = np.pi / 2
        assert abs(g[j] - 0.5 * (oracle_e(p + d) - oracle_e(p - d))) < 1e-5
    # the plain energy takes the fused expectation passes, the adjoint H|psi>
    assert abs(float(rt.energy(p, get_circuit)) - e) < 1e-6


def test_ucc_runtime_energy_and_grad(gpu):
    no, nv = 2, 2
    ns = no + nv
    n = 2 * ns
    h, g2 = chem.synthetic_integrals(ns, seed=7)
    terms = chem.molecular_hamiltonian(h, g2, e_core=0.3)
    # the same operator in OpenFermion form on spin-orbital indices
    of = {}
    for c, fac in terms:
        key = tuple(sorted((n - 1 - q, P) for P, q in fac))
        of[key] = of.get(key, 0.0) + c
    ex_ops, ids = chem.uccsd_excitations(no, nv)
    rt = R.UCCNumericRuntime(n, (no, no), of, ex_ops=ex_ops, param_ids=ids)
    th = np.random.default_rng(3).normal(0, 0.1, size=rt.n_params)

    def oracle_e(t):
        c = chem.ucc_circuit(t, n, (no, no), ex_ops, ids)
        psi, _ = O.simulate(n, c.ops)
        return O.pauli_sum_expectation(psi, terms, n)

    e, g = rt.energy_and_grad(th)
    assert abs(e - oracle_e(th)) < 1e-6
    for j in (0, rt.n_params // 2, rt.n_params - 1):
        d = np.zeros_like(th)
        d[j] = 1e-4
        fd = (oracle_e(th + d) - oracle_e(th - d)) / 2e-4
        assert abs(g[j] - fd) < 1e-5
    e2, gfd = rt.energy_and_grad(th, gradient_method="fd")
    assert abs(e2 - e) < 1e-12
    np.testing.assert_allclose(gfd[:4], g[:4], atol=2e-3)
    with pytest.raises(NotImplementedError):
        R.UCCNumericRuntime(n, (no, no), of, ex_ops=ex_ops, param_ids=ids, mode="qubit")
