"""The density-matrix oracle pinned by the statevector oracle, channel
identities and closed forms (CPU)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import density as DM
from oracle import statevector as O
from tests.conftest import load_known_answers


def _random_ops(rng, n, depth, names=("h", "rx", "ry", "rz", "cx", "cz", "x", "s", "sdg", "cry")):
    ops = []
    for _ in range(depth):
        name = names[int(rng.integers(len(names)))]
        if name in ("cx", "cz", "cry") and n < 2:
            name = "h"
        if name in ("cx", "cz", "cry"):
            a, b = (int(v) for v in rng.choice(n, 2, replace=False))
            ops.append((name, a, b) if name != "cry" else (name, a, b, float(rng.uniform(0, 6))))
        elif name in ("rx", "ry", "rz"):
            ops.append((name, int(rng.integers(n)), float(rng.uniform(0, 6))))
        else:
            ops.append((name, int(rng.integers(n))))
    return ops


@pytest.mark.parametrize("n", [1, 3, 4])
def test_pure_states_match_statevector(n):
    rng = np.random.default_rng(n)
    ops = _random_ops(rng, n, 30)
    rho, _ = DM.run(n, ops)
    psi, _ = O.simulate(n, ops, reference_dispatch=False)
    np.testing.assert_allclose(rho, np.outer(psi, psi.conj()), atol=1e-12)
    for q in range(n):
        assert abs(DM.exp_z(rho, q, n) - O.expect_z(psi, q, n)) < 1e-12


@pytest.mark.parametrize("kops", [DM.depolarizing(0.2), DM.amplitude_damping(0.3),
                                  DM.phase_damping(0.4), DM.pauli_channel(0.1, 0.05, 0.2)])
def test_channels_are_trace_preserving(kops):
    np.testing.assert_allclose(sum(K.conj().T @ K for K in kops), np.eye(2), atol=1e-12)
    rng = np.random.default_rng(0)
    rho, _ = DM.run(3, _random_ops(rng, 3, 12))
    out = DM.apply_kraus(rho, kops, 1, 3)
    assert abs(np.trace(out) - 1) < 1e-12
    np.testing.assert_allclose(out, out.conj().T, atol=1e-12)


def test_closed_forms_and_known_answer():
    rho, _ = DM.run(1, [("x", 0)], noise={"type": "amplitude_damping", "gamma": 0.3})
    assert abs(DM.exp_z(rho, 0, 1) - (1 - 2 * (1 - 0.3))) < 1e-12
    rho = np.zeros((2, 2), dtype=complex)
    rho[0, 0] = 1
    assert abs(DM.exp_z(DM.apply_kraus(rho, DM.depolarizing(0.3), 0, 1), 0, 1)
               - (1 - 4 * 0.3 / 3)) < 1e-12
    # the reference's noisy known answer (the statevector engine attenuates <Z>
    # by 1 - 4p/3; the density matrix applies the channel after the gate)
    case = {k["name"]: k for k in load_known_answers()}["rx_pi_z_depolarizing"]
    rho, meas = DM.run(case["n"], [tuple(o) for o in case["ops"]], noise=case["noise"])
    assert meas == [0]
    assert abs(DM.exp_z(rho, 0, 1) - case["run_expectations"]["Z0"]) < case["atol"]


def test_projection_renormalises():
    rho, _ = DM.run(2, [("h", 0), ("cx", 0, 1), ("project_z", 0, 1)])
    np.testing.assert_allclose(np.diag(rho).real, [0, 0, 0, 1], atol=1e-12)
