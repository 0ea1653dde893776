"""UCCSD energy and gradient (JW Pauli rotations, synthetic integrals) vs the
complex128 oracle (GPU).

The ansatz and Hamiltonian come from tyxonq_amd.chem (pinned against
brute-force fermionic matrices in tests/test_chem.py).  The gradient with
respect to the shared excitation parameters flows from the adjoint sweep
through torch autograd; the oracle gradient is a central difference of the
complex128 energy (h = 1e-4, FP64).  Tolerances: |dE| <= 1e-6 and
|dgrad| <= 1e-5 absolute, for gradients of up to about 4 from the synthetic
integrals.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import statevector as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("no,nv", [(2, 2), (3, 3)])
def test_ucc_energy_and_grad(gpu, no, nv):
    import torch

    from tyxonq_amd import chem
    from tyxonq_amd.engine import StatevectorEngine

    ns = no + nv
    n = 2 * ns
    h, g = chem.synthetic_integrals(ns, seed=11)
    terms = chem.molecular_hamiltonian(h, g)
    ex, ids = chem.uccsd_excitations(no, nv)
    npar = max(ids) + 1
    th = np.random.default_rng(7).normal(0.0, 0.05, size=npar)

    tt = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    c = chem.ucc_circuit([tt[i] for i in range(npar)], n, (no, no), ex, ids)
    e = StatevectorEngine(backend_name="pytorch").energy(c, terms)
    e.backward()
    grad = tt.grad.numpy()

    def e_ref(p):
        psi, _ = O.simulate(n, chem.ucc_circuit(p, n, (no, no), ex, ids).ops)
        return O.pauli_sum_expectation(psi, terms, n)

    assert abs(float(e) - e_ref(th)) < 1e-6
    hstep = 1e-4
    sample = range(npar) if npar <= 24 else np.random.default_rng(1).choice(npar, 24, replace=False)
    for k in sample:
        d = np.zeros(npar)
        d[k] = hstep
        fd = (e_ref(th + d) - e_ref(th - d)) / (2 * hstep)
        assert abs(grad[k] - fd) < 1e-5, (k, grad[k], fd)


def test_ucc_direct_params_match_autograd(gpu):
    # the op array with shared parameter ids and angle coefficients gives the
    # gradient for the shared parameters directly (no autograd through angles)
    import torch

    from tyxonq_amd import _native as nat
    from tyxonq_amd import chem
    from tyxonq_amd.engine import StatevectorEngine
    from tyxonq_amd.pauli import PauliSum

    no = nv = 3
    n = 2 * (no + nv)
    h, g = chem.synthetic_integrals(no + nv, seed=4)
    terms = chem.molecular_hamiltonian(h, g)
    ex, ids = chem.uccsd_excitations(no, nv)
    ans = chem.UCCAnsatz(n, (no, no), ex, ids)
    th = np.random.default_rng(9).normal(0.0, 0.05, size=ans.nparams)
    x, z, c = PauliSum(n, terms).to_masks()
    ham = nat.Ham(n, x, z, c)
    plan = nat.Plan(n, ans.ops(th * 0.5), max_cost=100000)
    plan.update(ans.ops(th))  # same structure: selection reused, programs re-emitted
    psi = nat.new_state(n, 0)
    lam = torch.empty_like(psi)
    e, grad = plan.adjoint(ham, psi, lam)
    tt = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    circ = chem.ucc_circuit([tt[i] for i in range(ans.nparams)], n, (no, no), ex, ids)
    ea = StatevectorEngine(backend_name="pytorch").energy(circ, terms)
    ea.backward()
    assert abs(e - float(ea.detach())) < 1e-6
    np.testing.assert_allclose(grad, tt.grad.numpy(), atol=1e-5)
