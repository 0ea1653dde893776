"""Generated pass kernels (compiled per plan at run time) vs the oracle (GPU).

Plans are forced onto generated kernels at small sizes (jit=1 or
TQSV_JIT=1): random circuits over every step kind and tile size, arbitrary
matrices, parameter updates (including rotation-form changes), the dual
sweep of the adjoint gradient and read-only expectation passes.  Same
tolerances as the precompiled kernel's tests.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import statevector as O
from tests.gpu_util import (dense_apply_items, ops_from_items, rand_state, random_circuit_ops,
                            random_matrix_items, random_pauli_terms, to_gpu, to_np)

pytestmark = pytest.mark.gpu


def _plan_run(n, ops, v, **opts):
    from tyxonq_amd import _native as nat

    psi = to_gpu(v)
    nat.Plan(n, ops, **opts).run(psi)
    return to_np(psi)


def _lowered(n, ops_tuples):
    import tyxonq_amd as tq
    from tyxonq_amd.engine import lower

    L, _, _ = lower(tq.Circuit(n, ops=list(ops_tuples)), mode="run", reference_dispatch=False)
    assert len(L.segments) == 1
    return L.segments[0][1]


@pytest.mark.parametrize("n,opts", [(11, {}), (13, {"tile_bits": 13}), (14, {"tile_bits": 14}),
                                    (15, {"tile_bits": 12, "max_cost": 24}),
                                    (16, {"tile_bits": 11, "max_cost": 5000}),
                                    (16, {"min_low_bits": 3})])
def test_random_circuits_generated(gpu, n, opts):
    rng = np.random.default_rng(500 + n)
    ops = random_circuit_ops(rng, n, 150)
    v = rand_state(rng, n)
    want, _ = O.simulate(n, ops, mode="run", init=v, reference_dispatch=False)
    got = _plan_run(n, _lowered(n, ops), v, jit=1, **opts)
    np.testing.assert_allclose(got, want, atol=2e-5)


@pytest.mark.parametrize("seed", range(3))
def test_arbitrary_matrices_generated(gpu, seed):
    n = 13
    rng = np.random.default_rng(900 + seed)
    items = random_matrix_items(rng, n, 160)
    v = rand_state(rng, n)
    want = dense_apply_items(n, items, v)
    scale = np.abs(want).max()
    for opts in ({}, {"max_cost": 100000}):
        got = _plan_run(n, ops_from_items(n, items), v, jit=1, **opts)
        np.testing.assert_allclose(got / scale, want / scale, atol=5e-5)


def test_pauli_rotations_generated(gpu):
    n = 14
    rng = np.random.default_rng(11)
    ops = []
    for _ in range(40):
        w = int(rng.integers(1, 7))
        qs = rng.choice(n, w, replace=False)
        ops.append(("pauli_rot", tuple(("XYZ"[int(rng.integers(3))], int(q)) for q in qs),
                    float(rng.uniform(-3, 3))))
        ops.append(("ry", int(rng.integers(n)), float(rng.uniform(0, 6))))
    v = rand_state(rng, n)
    want, _ = O.simulate(n, ops, mode="run", init=v, reference_dispatch=False)
    np.testing.assert_allclose(_plan_run(n, _lowered(n, ops), v, jit=1), want, atol=2e-5)


def test_update_across_rotation_forms(gpu):
    # angles move across |t| = 1 and beyond; the plan keeps working (with the
    # kernels it compiled or new ones) and matches the oracle after each update
    from tyxonq_amd import _native as nat

    n = 13
    rng = np.random.default_rng(4)
    base = rng.uniform(0, 2 * np.pi, 3 * n)

    def ops_for(shift):
        ops = []
        for q in range(n):
            ops += [("ry", q, float(base[q] + shift)), ("rx", q, float(base[n + q] - shift))]
        for q in range(0, n - 1, 2):
            ops.append(("cz", q, q + 1))
        for q in range(n):
            ops.append(("ry", q, float(base[2 * n + q] + 2 * shift)))
        return ops

    v = rand_state(rng, n)
    plan = nat.Plan(n, _lowered(n, ops_for(0.0)), jit=1)
    for shift in (0.0, 0.7, 1.6, 2.9, 0.1):
        plan.update(_lowered(n, ops_for(shift)))
        psi = to_gpu(v)
        plan.run(psi)
        want, _ = O.simulate(n, ops_for(shift), mode="run", init=v, reference_dispatch=False)
        np.testing.assert_allclose(to_np(psi), want, atol=2e-5)


@pytest.mark.parametrize("env,n", [({}, 12), ({"TQSV_PIPE": "1"}, 13), ({"TQSV_PIPE": "0"}, 14),
                                   ({"TQSV_DUAL1": "0"}, 15)])
def test_adjoint_generated(gpu, monkeypatch, env, n):
    # TQSV_PIPE unset or 1: gate passes stage the next tile in registers; 0:
    # none does.  TQSV_DUAL1=0: dual passes keep psi and lambda in two tile
    # buffers instead of one.  Each size gets its own cached plan; setting the
    # variables after earlier compiles also covers the compiler namespace's
    # environment.
    import torch

    import tyxonq_amd as tq
    from tyxonq_amd.engine import StatevectorEngine

    monkeypatch.setenv("TQSV_JIT", "1")
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    layers = 3
    rng = np.random.default_rng(31)
    th = rng.normal(0, 0.5, size=2 * n * layers)
    terms = random_pauli_terms(rng, n, 25, maxw=3)

    def ops_of(p):
        ops = []
        k = 0
        for _ in range(layers):
            for q in range(n):
                ops.append(("ry", q, p[k]))
                k += 1
            for q in range(n):
                ops.append(("rz", q, p[k]))
                k += 1
            for q in range(n):
                ops.append(("cx", q, (q + 1) % n))
        return ops

    def f(p):
        psi, _ = O.simulate(n, ops_of(p))
        return O.pauli_sum_expectation(psi, terms, n)

    tt = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    c = tq.Circuit(n, ops=ops_of([tt[i] for i in range(len(th))]))
    e, g = StatevectorEngine(backend_name="pytorch").energy_and_grad(c, terms)
    assert abs(e - f(th)) < 1e-6
    np.testing.assert_allclose(g, O.parameter_shift_gradient(f, th), atol=1e-5)


@pytest.mark.parametrize("n,maxw", [(11, 4), (14, 3), (16, 6)])
def test_expval_generated(gpu, monkeypatch, n, maxw):
    from tyxonq_amd import _native as nat
    from tyxonq_amd.pauli import PauliSum

    monkeypatch.setenv("TQSV_JIT", "1")
    rng = np.random.default_rng(70 + n)
    v = rand_state(rng, n)
    terms = random_pauli_terms(rng, n, 80, maxw=maxw) + [(0.5, [])]
    x, z, cf = PauliSum(n, terms).to_masks()
    e, nrm, per = nat.Ham(n, x, z, cf).expval(to_gpu(v), per_term=True)
    e_ref, vals = O.pauli_sum_expectation(v, terms, n, per_term=True)
    np.testing.assert_allclose(per, vals, atol=2e-6)
    assert abs(e - e_ref) < 1e-5 * sum(abs(c) for c, _ in terms)
    assert abs(nrm - 1.0) < 2e-6
