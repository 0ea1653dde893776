"""Single-operation HIP kernels of libtqsv vs the CPU oracle (GPU)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import statevector as O
from tests.gpu_util import rand_state, rand_unitary, to_gpu, to_np

pytestmark = pytest.mark.gpu

TOL = 3e-6


@pytest.mark.parametrize("n", [1, 2, 5, 12])
def test_apply_1q_every_bit(gpu, n):
    from tyxonq_amd import _native as nat

    rng = np.random.default_rng(n)
    v = rand_state(rng, n)
    for q in range(n):
        U = rand_unitary(rng, 2)
        psi = to_gpu(v)
        nat.apply_1q(psi, n - 1 - q, U)
        np.testing.assert_allclose(to_np(psi), O.apply_1q(v, U, q, n), atol=TOL)


@pytest.mark.parametrize("n", [2, 3, 6, 11])
def test_apply_2q_pairs(gpu, n):
    from tyxonq_amd import _native as nat

    rng = np.random.default_rng(10 + n)
    v = rand_state(rng, n)
    pairs = [(a, b) for a in range(n) for b in range(n) if a != b][:12]
    for q0, q1 in pairs:
        U = rand_unitary(rng, 4)
        psi = to_gpu(v)
        nat.apply_2q(psi, n - 1 - q0, n - 1 - q1, U)
        np.testing.assert_allclose(to_np(psi), O.apply_2q(v, U, q0, q1, n), atol=TOL)
    psi = to_gpu(v)
    nat.apply_2q(psi, 0, 0, rand_unitary(rng, 4))  # equal bits: no-op, as in the reference
    np.testing.assert_allclose(to_np(psi), v, atol=1e-7)


def test_apply_ctrl_and_diag(gpu):
    from tyxonq_amd import _native as nat

    n = 7
    rng = np.random.default_rng(3)
    v = rand_state(rng, n)
    U = rand_unitary(rng, 2)
    # control qubit 1 (bit 5) = 1, target qubit 4 (bit 2)
    psi = to_gpu(v)
    nat.apply_ctrl(psi, 1 << 5, 1 << 5, 2, U)
    want = O.apply_2q(v, np.block([[np.eye(2), np.zeros((2, 2))], [np.zeros((2, 2)), U]]), 1, 4, n)
    np.testing.assert_allclose(to_np(psi), want, atol=TOL)
    # control value 0
    psi = to_gpu(v)
    nat.apply_ctrl(psi, 1 << 5, 0, 2, U)
    want = O.apply_2q(v, np.block([[U, np.zeros((2, 2))], [np.zeros((2, 2)), np.eye(2)]]), 1, 4, n)
    np.testing.assert_allclose(to_np(psi), want, atol=TOL)
    # 3-bit diagonal on bits (6, 0, 3) -> qubits (0, 6, 3)
    ph = np.exp(1j * rng.uniform(0, 2 * np.pi, size=8))
    psi = to_gpu(v)
    nat.apply_diag(psi, [6, 0, 3], ph)
    idx = np.arange(1 << n)
    k = (((idx >> 6) & 1) << 2) | (((idx >> 0) & 1) << 1) | ((idx >> 3) & 1)
    np.testing.assert_allclose(to_np(psi), v * ph[k], atol=TOL)


@pytest.mark.parametrize("paulis", [[("Z", 0)], [("X", 1)], [("Y", 2), ("Z", 0)],
                                    [("X", 0), ("Y", 3), ("Z", 5), ("X", 6)],
                                    [("Y", q) for q in range(7)]])
def test_pauli_rot(gpu, paulis):
    from tyxonq_amd import _native as nat

    n = 7
    rng = np.random.default_rng(len(paulis))
    v = rand_state(rng, n)
    th = 0.917
    xm = zm = 0
    for P, q in paulis:
        b = 1 << (n - 1 - q)
        xm |= b if P in "XY" else 0
        zm |= b if P in "ZY" else 0
    psi = to_gpu(v)
    nat.pauli_rot(psi, xm, zm, th)
    np.testing.assert_allclose(to_np(psi), O.pauli_rotation(v, paulis, th, n), atol=TOL)


def test_reductions_and_projection(gpu):
    from tyxonq_amd import _native as nat

    n = 13
    rng = np.random.default_rng(9)
    v = rand_state(rng, n) * 1.3
    psi = to_gpu(v)
    assert nat.norm2(psi) == pytest.approx(np.vdot(v, v).real, rel=1e-6)
    for q in (0, 5, 12):
        assert nat.expect_z(psi, n - 1 - q) == pytest.approx(O.expect_z(v, q, n) , abs=2e-6)
    u = rand_state(rng, n)
    d = nat.dot(to_gpu(u), psi)
    assert abs(d - np.vdot(u, v)) < 2e-6
    # Pauli-weighted dot: sum (-1)^{|i&z|} conj(u_i) v_{i^x}
    x, z = 0b101, 0b110
    idx = np.arange(1 << n)
    sg = 1 - 2 * (np.array([bin(i & z).count("1") for i in idx]) & 1)
    assert abs(nat.dot(to_gpu(u), psi, x, z) - np.sum(sg * np.conj(u) * v[idx ^ x])) < 2e-6
    p = nat.probs(psi).cpu().numpy()
    np.testing.assert_allclose(p, np.abs(v) ** 2, rtol=1e-5, atol=1e-9)
    psi = to_gpu(v)
    kept = nat.project_z(psi, n - 1 - 4, 1)
    want = O.project_z(v, 4, 1, n)
    np.testing.assert_allclose(to_np(psi), want, atol=TOL)
    assert kept == pytest.approx(np.sum(np.abs(v.reshape((2,) * n)[(slice(None),) * 4 + (1,)]) ** 2),
                                 rel=1e-5)


@pytest.mark.parametrize("n,index", [(1, 1), (4, 9), (20, 12345)])
def test_init_basis(gpu, n, index):
    from tyxonq_amd import _native as nat

    psi = nat.new_state(n, index)
    a = to_np(psi)
    assert a[index] == 1 and np.count_nonzero(a) == 1
