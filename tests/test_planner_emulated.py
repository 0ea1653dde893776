"""The pass planner and program format, executed by the host emulator of the
tile-pass kernel (tqsv_plan_emulate) and compared with the oracle (CPU only).

The emulator interprets the same program words with the same index
arithmetic as the device kernel, so these tests catch scheduling, ordering
and encoding errors without a GPU.
"""
from __future__ import annotations

import numpy as np
import pytest

import tyxonq_amd as tq
from oracle import statevector as O
from tests.gpu_util import rand_state, random_circuit_ops
from tyxonq_amd import _native as nat
from tyxonq_amd.engine import lower


def _emulated(n, ops, v, **opts):
    L, _, _ = lower(tq.Circuit(n, ops=list(ops)), mode="run", reference_dispatch=False)
    got = v.astype(np.complex64)
    for seg in L.segments:
        got = nat.emulate(n, seg[1], got, **opts)
    return got


CASES = [(11, {}), (12, {}), (13, {"tile_bits": 13}), (14, {"tile_bits": 14}),
         (15, {"tile_bits": 12, "max_cost": 24}), (16, {"tile_bits": 11, "max_cost": 8}),
         (16, {"tile_bits": 13, "max_cost": 5000, "min_low_bits": 1}),
         (16, {"tile_bits": 14, "min_low_bits": 3}), (17, {"min_low_bits": 8}),
         (14, {"fuse": False})]


@pytest.mark.parametrize("n,opts", CASES, ids=[f"{n}-{sorted(o.items())}" for n, o in CASES])
def test_random_circuits_emulated(n, opts):
    for seed in range(2):
        rng = np.random.default_rng(1000 * n + seed)
        ops = random_circuit_ops(rng, n, 150)
        v = rand_state(rng, n)
        want, _ = O.simulate(n, ops, mode="run", init=v, reference_dispatch=False)
        np.testing.assert_allclose(_emulated(n, ops, v, **opts), want, atol=2e-5)


def test_commutation_with_converted_diagonals():
    # a CZ written as a 4x4 matrix becomes a diagonal gate; it must not move
    # ahead of the controlled-Y that shares its bit
    n = 11
    ops = [("cy", 2, 6), ("cry", 2, 9, 1.4), ("rxx", 3, 5, 0.51), ("rxx", 0, 5, 2.9),
           ("cy", 4, 8), ("cz", 8, 1)]
    v = rand_state(np.random.default_rng(1), n)
    want, _ = O.simulate(n, ops, mode="run", init=v, reference_dispatch=False)
    np.testing.assert_allclose(_emulated(n, ops, v), want, atol=1e-6)


@pytest.mark.parametrize("width", [1, 3, 5, 7])
def test_pauli_rotations_emulated(width):
    n = 13
    rng = np.random.default_rng(width)
    ops = []
    for _ in range(25):
        qs = rng.choice(n, width, replace=False)
        ops.append(("pauli_rot", tuple(("XYZ"[int(rng.integers(3))], int(q)) for q in qs),
                    float(rng.uniform(-3, 3))))
        ops.append(("rx", int(rng.integers(n)), 0.3))
    v = rand_state(rng, n)
    want, _ = O.simulate(n, ops, mode="run", init=v, reference_dispatch=False)
    np.testing.assert_allclose(_emulated(n, ops, v), want, atol=2e-5)


def test_brickwork_emulated_and_plan_limits():
    from tyxonq_amd import bench_workloads as W

    n = 18
    c = W.brickwork_circuit(n, 8, seed=5)
    v = np.zeros(1 << n, dtype=complex)
    v[0] = 1
    want, _ = O.simulate(n, c.ops)
    L, _, _ = lower(c)
    for opts in ({}, {"max_cost": 100000}, {"tile_bits": 14}):
        got = nat.emulate(n, L.segments[0][1], v, **opts)
        np.testing.assert_allclose(got, want, atol=2e-5)
    # the 30-qubit plan stays within the kernel's tile limits for every option
    c30 = W.brickwork_circuit(30, 20)
    L30, _, _ = lower(c30)
    for opts in ({}, {"tile_bits": 14}, {"min_low_bits": 1}, {"max_cost": 100000}):
        info = nat.inspect(30, L30.segments[0][1], **opts)
        assert info["passes"] >= 1
