"""Density-matrix engine on the GPU against the complex128 oracle."""
from __future__ import annotations

import numpy as np
import pytest

import tyxonq_amd as tq
from oracle import density as DM
from oracle import sampling as OS
from tests.test_density_oracle import _random_ops

pytestmark = pytest.mark.gpu

NOISES = [None, {"type": "depolarizing", "p": 0.05}, {"type": "amplitude_damping", "gamma": 0.1},
          {"type": "phase_damping", "lambda": 0.2}, {"type": "pauli", "px": 0.03, "py": 0.01,
                                                      "pz": 0.05}]


@pytest.mark.parametrize("k", range(len(NOISES)))
def test_run_expectations_and_expval(gpu, k):
    from tyxonq_amd.density import DensityMatrixEngine

    noise = NOISES[k]
    n = 3 + k % 3
    rng = np.random.default_rng(70 + k)
    ops = _random_ops(rng, n, 30) + [("measure_z", q) for q in range(n)]
    c = tq.Circuit(n, ops=list(ops))
    kw = {"use_noise": True, "noise": noise} if noise else {}
    eng = DensityMatrixEngine()
    out = eng.run(c, shots=0, **kw)
    rho, _ = DM.run(n, ops, noise=noise)
    for q in range(n):
        assert abs(out["expectations"][f"Z{q}"] - DM.exp_z(rho, q, n)) < 2e-6
    got = eng.density_matrix(c, **kw).cpu().numpy()
    np.testing.assert_allclose(got, rho, atol=3e-6)
    terms = [(0.7, [("X", 0), ("Y", 1)]), (-0.4, [("Z", 0)]), (1.3, [("Y", n - 1), ("X", 1)]),
             (0.25, [])]
    assert abs(eng.expval(c, terms, **kw) - DM.pauli_sum_expectation(rho, terms, n)) < 2e-6


def test_projection_reset_kraus_and_dispatch(gpu):
    from tyxonq_amd.density import DensityMatrixEngine

    n = 3
    c = tq.Circuit(n).h(0).cx(0, 1).ry(2, 0.8)
    c.ops += [("project_z", 0, 1), ("t", 2), ("reset", 2), ("measure_z", 0), ("measure_z", 2)]
    # the reference's density-matrix run skips t (not in its gate list)
    rho, _ = DM.run(n, c.ops)
    eng = DensityMatrixEngine()
    np.testing.assert_allclose(eng.density_matrix(c).cpu().numpy(), rho, atol=3e-6)
    out = eng.run(c, shots=0)
    assert abs(out["expectations"]["Z0"] + 1) < 1e-6 and abs(out["expectations"]["Z2"] - 1) < 1e-6
    # a circuit-defined channel
    c2 = tq.Circuit(2).x(0)
    c2._kraus_cache = {"ad": DM.amplitude_damping(0.25)}
    c2.ops += [("kraus", 0, "ad"), ("measure_z", 0)]
    r2, _ = DM.run(2, c2.ops, kraus_cache=c2._kraus_cache)
    assert abs(eng.run(c2, shots=0)["expectations"]["Z0"] - DM.exp_z(r2, 0, 2)) < 1e-6


def test_counts_from_the_diagonal(gpu):
    from tyxonq_amd import _native as nat
    from tyxonq_amd.density import DensityMatrixEngine

    n = 4
    rng = np.random.default_rng(5)
    ops = _random_ops(rng, n, 25)
    kc = {"dep": DM.depolarizing(0.2), "ad": DM.amplitude_damping(0.3)}
    ops += [("kraus", 0, "dep"), ("kraus", 2, "ad")] + [("measure_z", q) for q in range(n)]
    c = tq.Circuit(n, ops=list(ops))
    c._kraus_cache = kc
    eng = DensityMatrixEngine()
    counts = eng.run(c, shots=20000, seed=3)["result"]
    assert sum(counts.values()) == 20000
    # without use_noise the shots are the GPU draw: the oracle's sampler on the
    # same diagonal amplitudes gives the same counts
    amp = nat.dm_diag_amplitudes(eng.last_vec, n).cpu().numpy()
    widx, wcnt = OS.sample_counts(amp, 20000, 3)
    assert counts == {format(int(i), f"0{n}b"): int(k) for i, k in zip(widx, wcnt)}
    rho, _ = DM.run(n, ops, kraus_cache=kc)
    p = np.real(np.diag(rho))
    np.testing.assert_allclose(np.abs(amp) ** 2, p, atol=3e-6)


def test_noisy_counts_follow_the_mixed_diagonal(gpu):
    from tyxonq_amd.density import DensityMatrixEngine

    # with use_noise the diagonal is mixed once more before a host draw, as in
    # density_matrix/engine.py:108-127
    n = 3
    rng = np.random.default_rng(9)
    ops = _random_ops(rng, n, 20) + [("measure_z", q) for q in range(n)]
    noise = {"type": "depolarizing", "p": 0.1}
    shots = 40000
    counts = DensityMatrixEngine().run(tq.Circuit(n, ops=list(ops)), shots=shots, use_noise=True,
                                       noise=noise)["result"]
    rho, _ = DM.run(n, ops, noise=noise)
    a = 4 * 0.1 / 3
    p = (1 - a) * np.clip(np.real(np.diag(rho)), 0, None) + a / (1 << n)
    p /= p.sum()
    got = np.array([counts.get(format(i, f"0{n}b"), 0) for i in range(1 << n)])
    assert got.sum() == shots
    sd = np.sqrt(shots * p * (1 - p))
    assert np.all(np.abs(got - shots * p) <= 5 * sd + 1)


def test_driver_routes_density_matrix(gpu):
    c = tq.Circuit(2).h(0).cx(0, 1).measure_z(0).measure_z(1)
    c.device(provider="simulator", device="density_matrix")
    r = c.run(shots=0)[0]
    # the engine payload sits under result_meta, as for the statevector device
    assert abs(r["result_meta"]["expectations"]["Z0"]) < 1e-6
    counts = c.run(shots=4000)[0]["result"]
    assert sum(counts.values()) == 4000 and set(counts) <= {"00", "11"}


def test_seven_qubit_noisy_circuit(gpu):
    from tyxonq_amd.density import DensityMatrixEngine

    n = 7
    rng = np.random.default_rng(77)
    ops = _random_ops(rng, n, 60)
    kc = {"ad": DM.amplitude_damping(0.15), "pd": DM.phase_damping(0.3)}
    ops += [("kraus", 3, "ad"), ("kraus", 6, "pd")] + [("measure_z", q) for q in range(n)]
    noise = {"type": "pauli", "px": 0.02, "py": 0.01, "pz": 0.03}
    c = tq.Circuit(n, ops=list(ops))
    c._kraus_cache = kc
    eng = DensityMatrixEngine()
    rho, _ = DM.run(n, ops, noise=noise, kraus_cache=kc)
    got = eng.density_matrix(c, use_noise=True, noise=noise).cpu().numpy()
    np.testing.assert_allclose(got, rho, atol=5e-6)
    out = eng.run(c, shots=0, use_noise=True, noise=noise)
    for q in range(n):
        assert abs(out["expectations"][f"Z{q}"] - DM.exp_z(rho, q, n)) < 5e-6
    terms = [(0.5, [("X", 0), ("X", 6)]), (-1.2, [("Y", 2), ("Z", 3), ("X", 5)]), (0.3, [("Z", 4)])]
    assert abs(eng.expval(c, terms, use_noise=True, noise=noise)
               - DM.pauli_sum_expectation(rho, terms, n)) < 5e-6
