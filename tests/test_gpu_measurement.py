"""Counts-based Pauli-sum energies and shared-prefix driver batches on the GPU,
against the oracle's statevector and its restatement of the GPU sampler."""
from __future__ import annotations

import numpy as np
import pytest

import tyxonq_amd as tq
from oracle import dense as D
from oracle import sampling as OS
from oracle import statevector as O
from tyxonq_amd import measurement as M
from tyxonq_amd.postprocessing import term_expectation_from_counts

pytestmark = pytest.mark.gpu


def _circuit(rng, n, layers):
    c = tq.Circuit(n)
    for _ in range(layers):
        for q in range(n):
            c.ry(q, float(rng.uniform(0, 2 * np.pi))).rz(q, float(rng.uniform(0, 2 * np.pi)))
        for q in range(n - 1):
            c.cx(q, q + 1)
    return c


def _hamiltonian(rng, n, k):
    ham = [(0.7, [])]
    for _ in range(k):
        qs = rng.choice(n, size=int(rng.integers(1, 4)), replace=False)
        ham.append((float(rng.normal()), [(str(rng.choice(list("XYZ"))), int(q)) for q in qs]))
    return ham


@pytest.mark.parametrize("n,shots", [(6, 20000), (11, 50000)])
def test_grouped_energy_against_oracle(gpu, n, shots):
    rng = np.random.default_rng(n)
    c = _circuit(rng, n, 3)
    ham = _hamiltonian(rng, n, 12)
    res = M.grouped_pauli_energy(c, ham, shots, seed=17)
    psi_ref, _ = O.simulate(n, c.ops)
    psi_gpu = tq.StatevectorEngine(reference_dispatch=False).state_tensor(c).cpu().numpy()
    assert np.allclose(psi_gpu, psi_ref, atol=2e-6)
    const, groups = M.group_hamiltonian_pauli_terms(ham, n)
    assert len(res["groups"]) == len(groups) and res["identity_const"] == const
    energy = const
    for k, g in enumerate(res["groups"]):
        # the rotated copy is the oracle's rotation of the oracle state
        phi = M.rotate_to_bases(tq.StatevectorEngine(reference_dispatch=False).state_tensor(c),
                                g["bases"]).cpu().numpy()
        want = psi_ref
        for q, P in enumerate(g["bases"]):
            if P == "X":
                want = O.apply_1q(want, O.gate_h(), q, n)
            elif P == "Y":
                want = O.apply_1q(O.apply_1q(want, O.gate_sd(), q, n), O.gate_h(), q, n)
        assert np.allclose(phi, want, atol=3e-6)
        # the group's counts are the sampler's draw on that state with the group's seed
        widx, wcnt = OS.sample_counts(phi, shots, M.group_seed(17, k))
        np.testing.assert_array_equal(g["index"], widx)
        np.testing.assert_array_equal(g["count"], wcnt)
        counts = {format(int(i), f"0{n}b"): int(v) for i, v in zip(widx, wcnt)}
        for (term, coeff), ev in zip(g["items"], g["expvals"]):
            assert ev == pytest.approx(term_expectation_from_counts(counts, [q for q, _ in term]),
                                       abs=1e-12)
            energy += coeff * ev
    assert res["energy"] == pytest.approx(energy, abs=1e-10)
    # and within sampling error of the exact energy
    exact = float(np.real(np.vdot(psi_ref, D.pauli_sum_dense(ham, n) @ psi_ref)))
    sd = np.sqrt(sum(len(items) * sum(cf * cf for _, cf in items)
                     for items in groups.values()) / shots)
    assert abs(res["energy"] - exact) < 5 * sd


def test_driver_batch_shares_prefixes(gpu):
    from tyxonq_amd.engine import StatevectorEngine

    n = 10
    rng = np.random.default_rng(3)
    base = _circuit(rng, n, 4)
    rots = [[], [("h", 0), ("h", 3)], [("sdg", 2), ("h", 2)], [("h", 9)]]
    batch = [base.extended(r + [("measure_z", q) for q in range(n)]) for r in rots]
    other = _circuit(rng, n, 2).extended([("measure_z", q) for q in range(n)])
    batch = batch[:2] + [other] + batch[2:]
    eng = StatevectorEngine()
    outs = eng.run_many(batch, [4000] * len(batch), seed=5)
    for i, c in enumerate(batch):
        seed_i = int(np.random.SeedSequence([5, i]).generate_state(1, dtype=np.uint64)[0])
        ref = StatevectorEngine().run(c, shots=4000, seed=seed_i)["result"]
        got = outs[i]["result"]
        assert sum(got.values()) == 4000
        # same distribution; the shared-prefix state differs from a separate run
        # only by complex64 rounding
        psi, _ = O.simulate(n, c.ops, mode="run")
        p = np.abs(psi) ** 2
        for cnts in (got, ref):
            v = np.array([cnts.get(format(j, f"0{n}b"), 0) for j in range(1 << n)])
            assert np.all(np.abs(v - 4000 * p) <= 5 * np.sqrt(4000 * p * (1 - p)) + 2)
    # through the driver: a list of circuits returns one result per circuit
    tasks = tq.devices.run(provider="simulator", device="statevector", circuit=batch, shots=1000)
    res = [t.get_result() for t in tasks]
    assert len(res) == len(batch) and all(sum(r["result"].values()) == 1000 for r in res)
