"""Fused tile passes (libtqsv plans) vs the CPU oracle (GPU).

Covers every step kind of the precompiled pass kernel: rotations with
deferred scale in all four forms, one-sided phases, sign flips, slot tables,
per-thread factors and rescaling of the pending diagonal, general and
controlled 2x2, 4x4 on slot pairs, 2-bit diagonals that are not phases, Pauli
rotations with 1..5 flip bits (and >5 through the single-op path), at tile
sizes 2^11..2^14, with small and large per-pass budgets.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import statevector as O
from tests.gpu_util import rand_state, random_circuit_ops, to_gpu, to_np

pytestmark = pytest.mark.gpu


def _run_plan(n, ops_tuples, init, **opts):
    import tyxonq_amd as tq
    from tyxonq_amd import _native as nat
    from tyxonq_amd.engine import lower

    c = tq.Circuit(n, ops=list(ops_tuples))
    L, _, _ = lower(c, mode="run", reference_dispatch=False)
    psi = to_gpu(init)
    for seg in L.segments:
        plan = nat.Plan(n, seg[1], **opts)
        plan.run(psi)
    return to_np(psi), L


def _check(n, ops, seed, atol=2e-5, **opts):
    rng = np.random.default_rng(seed)
    v = rand_state(rng, n)
    got, _ = _run_plan(n, ops, v, **opts)
    want, _ = O.simulate(n, ops, mode="run", init=v, reference_dispatch=False)
    np.testing.assert_allclose(got, want, atol=atol)


@pytest.mark.parametrize("n", [3, 8, 10])
def test_small_states_single_op_path(gpu, n):
    rng = np.random.default_rng(n)
    _check(n, random_circuit_ops(rng, n, 60), seed=n)


@pytest.mark.parametrize("n,tile", [(11, 0), (12, 0), (13, 13), (14, 13), (14, 14), (16, 11),
                                    (16, 12), (17, 13), (18, 14)])
def test_random_circuits_tile_passes(gpu, n, tile):
    rng = np.random.default_rng(100 + n + tile)
    ops = random_circuit_ops(rng, n, 150)
    _check(n, ops, seed=n, tile_bits=tile)


@pytest.mark.parametrize("max_cost", [8, 24, 5000])
def test_pass_budget_variants(gpu, max_cost):
    n = 15
    rng = np.random.default_rng(max_cost)
    ops = random_circuit_ops(rng, n, 200)
    _check(n, ops, seed=7, max_cost=max_cost)


def test_unfused_plan_matches(gpu):
    n = 13
    rng = np.random.default_rng(5)
    ops = random_circuit_ops(rng, n, 80)
    _check(n, ops, seed=5, fuse=False)


@pytest.mark.parametrize("min_low", [1, 3, 8])
def test_min_low_bits_variants(gpu, min_low):
    n = 16
    rng = np.random.default_rng(min_low)
    ops = random_circuit_ops(rng, n, 120)
    _check(n, ops, seed=3, min_low_bits=min_low)


def test_brickwork_20q(gpu):
    from tyxonq_amd import bench_workloads as W

    n = 20
    c = W.brickwork_circuit(n, 6, seed=99)
    v = np.zeros(1 << n, dtype=complex)
    v[0] = 1
    got, _ = _run_plan(n, c.ops, v)
    want, _ = O.simulate(n, c.ops)
    np.testing.assert_allclose(got, want, atol=2e-5)


@pytest.mark.parametrize("width", [1, 2, 3, 4, 5, 6, 8])
def test_pauli_rotations_fused(gpu, width):
    n = 14
    rng = np.random.default_rng(width)
    ops = []
    for _ in range(30):
        qs = rng.choice(n, width, replace=False)
        fac = tuple(("XYZ"[int(rng.integers(3))], int(q)) for q in qs)
        ops.append(("pauli_rot", fac, float(rng.uniform(-3, 3))))
        ops.append(("h", int(rng.integers(n))))
    _check(n, ops, seed=width)


def test_plan_update_new_angles(gpu):
    import tyxonq_amd as tq
    from tyxonq_amd import _native as nat
    from tyxonq_amd.engine import lower

    n = 14
    rng = np.random.default_rng(2)
    ops1 = random_circuit_ops(rng, n, 60, allow={"rx", "ry", "rz", "cx", "rzz"})
    ops2 = [tuple(o[:-1]) + (o[-1] + 0.5,) if isinstance(o[-1], float) else o for o in ops1]
    L1, _, _ = lower(tq.Circuit(n, ops=ops1), mode="run", reference_dispatch=False)
    L2, _, _ = lower(tq.Circuit(n, ops=ops2), mode="run", reference_dispatch=False)
    plan = nat.Plan(n, L1.segments[0][1])
    plan.update(L2.segments[0][1])
    v = rand_state(rng, n)
    psi = to_gpu(v)
    plan.run(psi)
    want, _ = O.simulate(n, ops2, mode="run", init=v, reference_dispatch=False)
    np.testing.assert_allclose(to_np(psi), want, atol=2e-5)
