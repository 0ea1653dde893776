"""Adjoint energy+gradient and the state VJP vs the CPU oracle (GPU).

Checks the hand-written reverse sweep against three independent gradients
of the complex128 oracle: the parameter-shift rule, central finite
differences (h = 1e-3) and PyTorch-CPU autograd through the reference's
per-gate contractions.  Tolerance: max|dgrad| <= 1e-5 (c64 state, FP64
reductions) and |dE| <= 1e-6 for these sizes.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import statevector as O
from tests.gpu_util import random_pauli_terms, to_gpu, to_np

pytestmark = pytest.mark.gpu


def _ops_with_params(n, layers, th):
    ops = []
    k = 0
    for _ in range(layers):
        for q in range(n):
            ops.append(("ry", q, th[k]))
            k += 1
        for q in range(n):
            ops.append(("rz", q, th[k]))
            k += 1
        for q in range(n):
            ops.append(("cx", q, (q + 1) % n))
    return ops


def _energy_oracle(n, layers, terms):
    def f(p):
        psi, _ = O.simulate(n, _ops_with_params(n, layers, p))
        return O.pauli_sum_expectation(psi, terms, n)

    return f


@pytest.mark.parametrize("n,layers", [(3, 2), (11, 2), (13, 2)])
def test_adjoint_vs_oracle_gradients(gpu, n, layers):
    import torch

    import tyxonq_amd as tq
    from tyxonq_amd.engine import StatevectorEngine

    rng = np.random.default_rng(n)
    th = rng.normal(0, 0.5, size=2 * n * layers)
    terms = random_pauli_terms(rng, n, 30, maxw=3)
    tt = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    c = tq.Circuit(n, ops=_ops_with_params(n, layers, [tt[i] for i in range(len(th))]))
    e, g = StatevectorEngine(backend_name="pytorch").energy_and_grad(c, terms)
    f = _energy_oracle(n, layers, terms)
    e_ref = f(th)
    assert abs(e - e_ref) < 1e-6
    ps = O.parameter_shift_gradient(f, th)
    np.testing.assert_allclose(g, ps, atol=1e-5)
    idx = list(range(0, len(th), max(1, len(th) // 6)))
    fd = O.central_fd_gradient(f, th, h=1e-3, indices=idx)
    np.testing.assert_allclose(g[idx], fd, atol=1e-5)
    e_ag, ag = O.torch_autograd_energy(n, lambda t: _ops_with_params(n, layers, t), terms, th)
    np.testing.assert_allclose(g, ag, atol=1e-5)


def test_adjoint_all_rotation_kinds(gpu):
    import torch

    import tyxonq_amd as tq

    n = 12
    rng = np.random.default_rng(21)
    vals = rng.normal(0, 1.0, size=8)
    t = torch.tensor(vals, requires_grad=True)

    def build(p):
        return [("h", 0), ("h", 3), ("rx", 0, p[0]), ("ry", 5, p[1]), ("rz", 3, p[2]),
                ("rxx", 0, 5, p[3]), ("ryy", 3, 7, p[4]), ("rzz", 1, 11, p[5]),
                ("cx", 5, 11), ("pauli_rot", (("X", 2), ("Y", 6), ("Z", 9)), p[6]),
                ("pauli_rot", (("Y", 0), ("X", 4), ("X", 8), ("Y", 10)), p[7])]

    terms = random_pauli_terms(rng, n, 25, maxw=4)
    c = tq.Circuit(n, ops=build([t[i] for i in range(8)]))
    tq.set_backend("pytorch")
    try:
        E = c.expectation(terms)
        E.backward()
    finally:
        tq.set_backend("numpy")

    def f(p):
        psi, _ = O.simulate(n, build(p))
        return O.pauli_sum_expectation(psi, terms, n)

    assert abs(float(E) - f(vals)) < 1e-6
    np.testing.assert_allclose(t.grad.numpy(), O.parameter_shift_gradient(f, vals), atol=1e-5)


def test_state_vjp_through_torch(gpu):
    import torch

    import tyxonq_amd as tq

    n = 11
    rng = np.random.default_rng(5)
    vals = rng.normal(0, 0.8, size=2 * n)
    target = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    t = torch.tensor(vals, requires_grad=True)
    tq.set_backend("pytorch")
    try:
        c = tq.Circuit(n, ops=_ops_with_params(n, 1, [t[i] for i in range(len(vals))]))
        psi = c.state()
        tgt = torch.tensor(target.astype(np.complex64)).cuda()
        loss = torch.abs(torch.vdot(tgt, psi)) ** 2 + torch.sum(psi.real[:50] * psi.imag[:50])
        loss.backward()
    finally:
        tq.set_backend("numpy")

    def f(p):
        v, _ = O.simulate(n, _ops_with_params(n, 1, p))
        return abs(np.vdot(target, v)) ** 2 + float(np.sum(v.real[:50] * v.imag[:50]))

    fd = O.central_fd_gradient(f, vals, h=1e-4)
    scale = max(1.0, np.max(np.abs(fd)))
    np.testing.assert_allclose(t.grad.numpy(), fd, atol=2e-4 * scale)


def test_known_answer_zero_gradient(gpu):
    import torch

    import tyxonq_amd as tq

    t = torch.tensor([0.0, 0.5, 1.0], dtype=torch.float64, requires_grad=True)
    tq.set_backend("pytorch")
    try:
        es = []
        for i in range(3):
            c = tq.Circuit(1).rx(0, t[i])
            es.append(c.expectation([(1.0, [("Z", 0)])]))
        torch.stack(es).mean().backward()
    finally:
        tq.set_backend("numpy")
    # test_circuit_state_autograd.py:154-189: theta = 0 has zero gradient
    assert abs(float(t.grad[0])) < 1e-6
    np.testing.assert_allclose(t.grad.numpy(), -np.sin([0.0, 0.5, 1.0]) / 3, atol=1e-6)
