"""Density matrices as 2n-qubit states (CPU): the doubled gate rows and the
channel superoperators, run by the host emulator of the tile passes, give
the oracle's rho."""
from __future__ import annotations

import numpy as np
import pytest

import tyxonq_amd as tq
from oracle import density as DM
from tests.test_density_oracle import _random_ops
from tyxonq_amd import _native as nat
from tyxonq_amd import density as D
from tyxonq_amd.engine import lower


def _emulated_rho(n, ops, noise=None):
    rows = []
    S = D.superoperator(D.channel_kraus(noise)) if noise else None
    for op in ops:
        L, _, _ = lower(tq.Circuit(n, ops=[op]), mode="run", reference_dispatch=False)
        for seg in L.segments:
            rows += D.double_rows(seg[1], n)
        if S is not None:
            for q in D._wires(op):
                rows.append(D._superop_row(n - 1 - q, n, S))
    vec = np.zeros(1 << (2 * n), dtype=np.complex64)
    vec[0] = 1
    out = nat.emulate(2 * n, np.array(rows, dtype=nat.OP_DTYPE), vec)
    return out.reshape(1 << n, 1 << n)


@pytest.mark.parametrize("n,noise", [(2, None), (3, {"type": "depolarizing", "p": 0.1}),
                                     (3, {"type": "amplitude_damping", "gamma": 0.2}),
                                     (4, {"type": "phase_damping", "lambda": 0.15}),
                                     (4, {"type": "pauli", "px": 0.05, "py": 0.02, "pz": 0.1})])
def test_vectorised_density_matches_oracle(n, noise):
    rng = np.random.default_rng(40 + n)
    ops = _random_ops(rng, n, 25)
    want, _ = DM.run(n, ops, noise=noise)
    got = _emulated_rho(n, ops, noise)
    np.testing.assert_allclose(got, want, atol=3e-6)


def test_pauli_rotation_conjugate_copy():
    n = 3
    ops = [("h", 0), ("pauli_rot", (("Y", 0), ("X", 2)), 0.7), ("pauli_rot", (("Z", 1), ("Y", 2)), -1.1)]
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1
    from oracle import statevector as O

    psi, _ = O.simulate(n, ops, reference_dispatch=False)
    np.testing.assert_allclose(_emulated_rho(n, ops), np.outer(psi, psi.conj()), atol=3e-6)


def test_superoperator_is_the_channel():
    K = D.channel_kraus({"type": "amplitude_damping", "gamma": 0.3})
    rho = np.array([[0.3, 0.2 - 0.1j], [0.2 + 0.1j, 0.7]])
    want = sum(k @ rho @ k.conj().T for k in K)
    got = (D.superoperator(K) @ rho.reshape(-1)).reshape(2, 2)
    np.testing.assert_allclose(got, want, atol=1e-12)
