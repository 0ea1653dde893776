"""Pins the complex128 C restatement of the oracle (CPU only).

oracle/sv128.c (through oracle/c128.py) computes the full-size golden values.
It is checked here against independent constructions: dense Kronecker
operators (oracle/dense.py) for every primitive, scipy's matrix exponential
for Pauli rotations, the NumPy oracle for whole circuits, and the
parameter-shift rule and central differences for the reverse-sweep gradient.
A swapped bit order, a wrong Y phase, a transposed 4x4 or a sign error in
the gradient fails one of them.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

from oracle import c128 as C
from oracle import dense
from oracle import statevector as O
from tests.gpu_util import rand_state, rand_unitary

N = 5


@pytest.fixture(scope="module", autouse=True)
def _lib():
    C.lib()
    C.set_threads(2)


def test_apply_1q_every_qubit_vs_dense():
    rng = np.random.default_rng(1)
    for q in range(N):
        v = rand_state(rng, N)
        U = rand_unitary(rng, 2)
        psi = v.copy()
        C.apply_1q(psi, U, q, N)
        np.testing.assert_allclose(psi, dense.op_1q(U, q, N) @ v, atol=1e-13)


def test_apply_2q_every_ordered_pair_vs_dense():
    rng = np.random.default_rng(2)
    for q0 in range(N):
        for q1 in range(N):
            v = rand_state(rng, N)
            U = rand_unitary(rng, 4)
            psi = v.copy()
            C.apply_2q(psi, U, q0, q1, N)
            want = v if q0 == q1 else dense.op_2q(U, q0, q1, N) @ v
            np.testing.assert_allclose(psi, want, atol=1e-13)


def _rand_pauli(rng, n):
    while True:
        fac = [("IXYZ"[int(rng.integers(4))], q) for q in range(n)]
        fac = [(P, q) for P, q in fac if P != "I"]
        if fac:
            return fac


def test_pauli_rot_dot_apply_vs_dense():
    rng = np.random.default_rng(3)
    for _ in range(25):
        fac = _rand_pauli(rng, N)
        P = dense.pauli_string(fac, N)
        x, z = C.masks(fac, N)
        v = rand_state(rng, N)
        a = float(rng.uniform(-3, 3))
        psi = v.copy()
        C.pauli_rot(psi, x, z, a, N)
        np.testing.assert_allclose(psi, scipy.linalg.expm(-0.5j * a * P) @ v, atol=1e-12)
        u = rand_state(rng, N)
        assert C.pauli_dot(u, v, x, z, N) == pytest.approx(np.vdot(u, P @ v), abs=1e-12)
        out = np.zeros_like(v)
        C.lib().sv_pauli_apply_acc(C._p(v), C._p(out), N, x, z, 0.7)
        np.testing.assert_allclose(out, 0.7 * (P @ v), atol=1e-13)
    assert C.norm2(v, N) == pytest.approx(float(np.vdot(v, v).real), abs=1e-13)


def test_ham_apply_and_term_values_vs_dense():
    rng = np.random.default_rng(4)
    terms = [(float(rng.normal()), _rand_pauli(rng, N)) for _ in range(12)]
    v = rand_state(rng, N)
    H = dense.pauli_sum_dense(terms, N)
    np.testing.assert_allclose(C.ham_apply(v, terms, N), H @ v, atol=1e-12)
    vals = C.term_values(v, terms, N)
    want = [float(np.real(np.vdot(v, dense.pauli_string(f, N) @ v))) for _, f in terms]
    np.testing.assert_allclose(vals, want, atol=1e-12)


def test_whole_circuits_vs_numpy_oracle():
    from tyxonq_amd import bench_workloads as W

    n = 8
    for c in (W.brickwork_circuit(n, 5, seed=3), W.ryrz_ring_ansatz(n, 3)[0],
              W.hea_ry_cx(n, 3)[0]):
        ops = [tuple(op) for op in c.ops]
        want, _ = O.simulate(n, ops)
        psi = C.run_gates(C.basis(n), C.gates_of(ops, n), n)
        np.testing.assert_allclose(psi, want, atol=1e-12)
    ops = [("h", 0), ("rxx", 1, 3, 0.4), ("ryy", 2, 0, 1.1), ("rzz", 3, 1, -0.3), ("iswap", 0, 2),
           ("swap", 1, 2), ("s", 3), ("sdg", 0), ("x", 1),
           ("pauli_rot", (("X", 0), ("Y", 2), ("Z", 3)), 0.8)]
    want, _ = O.simulate(4, ops)
    np.testing.assert_allclose(C.run_gates(C.basis(4), C.gates_of(ops, 4), 4), want, atol=1e-12)


def test_reverse_sweep_gradient_vs_parameter_shift_and_fd():
    from tyxonq_amd import bench_workloads as W

    n, layers = 6, 2
    th = np.random.default_rng(5).normal(0, 0.5, size=2 * n * layers)
    terms = W.heisenberg_ring_terms(n, 0.5)

    def energy(p):
        psi, _ = O.simulate(n, W.ryrz_ring_ansatz(n, layers, params=list(p))[0].ops)
        return O.pauli_sum_expectation(psi, terms, n)

    c, _ = W.ryrz_ring_ansatz(n, layers, params=list(th))
    gates = C.gates_of(c.ops, n, param_of=[(k, 1.0) for k in range(th.size)])
    e, g, nrm = C.energy_and_grad(n, gates, terms, th.size)
    assert e == pytest.approx(energy(th), abs=1e-12)
    assert nrm == pytest.approx(1.0, abs=1e-12)
    np.testing.assert_allclose(g, O.parameter_shift_gradient(energy, th), atol=1e-11)
    np.testing.assert_allclose(g[:5], O.central_fd_gradient(energy, th, h=1e-5,
                                                            indices=range(5)), atol=1e-8)


def test_reverse_sweep_shared_parameters_ucc():
    """Shared parameters with angle coefficients (UCC rotations, a = coef * theta_p)."""
    from tyxonq_amd import chem

    ns, no = 4, 2
    n = 2 * ns
    h, gint = chem.synthetic_integrals(ns, 7)
    terms = chem.molecular_hamiltonian(h, gint)
    ex, ids = chem.uccsd_excitations(no, ns - no)
    npar = max(ids) + 1
    th = np.random.default_rng(8).normal(0, 0.3, size=npar)

    def ops_of(p):
        return [tuple(op) for op in chem.ucc_circuit(list(p), n, (no, no), ex, ids).ops]

    def energy(p):
        psi, _ = O.simulate(n, ops_of(p))
        return O.pauli_sum_expectation(psi, terms, n)

    param_of = []
    for f_idx, pid in zip(ex, ids):
        for v, _, _ in chem.excitation_strings(f_idx, n):
            param_of.append((pid, -2.0 * v.imag))
    gates = C.gates_of(ops_of(th), n, param_of=param_of)
    e, g, _ = C.energy_and_grad(n, gates, terms, npar)
    assert e == pytest.approx(energy(th), abs=1e-11)
    np.testing.assert_allclose(g, O.central_fd_gradient(energy, th, h=1e-5), atol=1e-7)
