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
from tests.gpu_util import (dense_apply_items, ops_from_items, rand_state, random_circuit_ops,
                            random_matrix_items)
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


@pytest.mark.parametrize("seed", range(6))
def test_decomposed_rotations_and_pending_diagonals(seed):
    # arbitrary 1-qubit unitaries (incl. scaled ones), diagonal gates on any two
    # bits (sign-only and general phases), controlled and non-unitary matrices
    n = 12
    rng = np.random.default_rng(77 + seed)
    items = random_matrix_items(rng, n, 160)
    v = rand_state(rng, n)
    want = dense_apply_items(n, items, v)
    ops = ops_from_items(n, items)
    for opts in ({}, {"max_cost": 100000}, {"tile_bits": 11, "max_cost": 12}):
        got = nat.emulate(n, ops, v.astype(np.complex64), **opts)
        scale = np.abs(want).max()
        np.testing.assert_allclose(got / scale, want / scale, atol=5e-5)


def test_long_rotation_chain_rescales():
    # hundreds of deferred-scale rotations inside one pass must not overflow
    n = 12
    rng = np.random.default_rng(3)
    items = []
    for i in range(600):
        q = int(rng.integers(3))
        items.append(("m1", q, O.gate_rx(rng.uniform(1.2, 1.9)) if i % 2 else O.gate_ry(1.5), 0, 0))
        items.append(("m2", q, (q + 1) % 3, np.diag([1, 1, 1, -1])))
    v = rand_state(rng, n)
    want = dense_apply_items(n, items, v)
    got = nat.emulate(n, ops_from_items(n, items), v.astype(np.complex64), max_cost=100000)
    assert np.isfinite(got).all()
    np.testing.assert_allclose(got, want, atol=2e-4)
    assert abs(np.linalg.norm(got) - 1) < 1e-4


@pytest.mark.parametrize("seed", range(4))
def test_controlled_swaps_move_pending_phases(seed):
    # antidiagonal unit-modulus gates (X, Y, X times phases), uncontrolled or
    # with one control of either value, interleaved with pending diagonals
    # (1-bit phases, sign and general 2-bit phases) and rotations, so that the
    # pending factor is moved through the swaps in every configuration:
    # control in the registers, on a thread bit or outside the tile
    n = 13
    rng = np.random.default_rng(500 + seed)
    items = []
    for _ in range(220):
        r = rng.random()
        q = int(rng.integers(n))
        c = int(rng.integers(n - 1))
        c = c if c < q else c + 1
        if r < 0.4:
            a, b = np.exp(1j * rng.uniform(0, 2 * np.pi, 2)) if rng.random() < 0.4 else (1, 1)
            if rng.random() < 0.2:
                a, b = -1j, 1j
            kind = int(rng.integers(3))
            cm = 0 if kind == 0 else 1 << (n - 1 - c)
            cv = cm if kind == 1 else 0
            items.append(("m1", q, np.array([[0, a], [b, 0]]), cm, cv))
        elif r < 0.65:
            d = [1, 1, 1, -1] if rng.random() < 0.5 else np.exp(1j * rng.uniform(0, 2 * np.pi, 4))
            items.append(("m2", q, c, np.diag(d)))
        elif r < 0.8:
            items.append(("m1", q, np.diag([1, np.exp(1j * rng.uniform(0, 2 * np.pi))]), 0, 0))
        else:
            items.append(("m1", q, O.gate_ry(rng.uniform(0, 6)), 0, 0))
    v = rand_state(rng, n)
    want = dense_apply_items(n, items, v)
    ops = ops_from_items(n, items)
    for opts in ({}, {"max_cost": 100000}, {"tile_bits": 11, "max_cost": 6},
                 {"tile_bits": 12, "min_low_bits": 2}):
        got = nat.emulate(n, ops, v.astype(np.complex64), **opts)
        np.testing.assert_allclose(got, want, atol=5e-5)


def test_cx_ladder_ansatz_pass_count():
    # RY/RZ layers with CX ladders: the CX steps are register renamings and
    # the RZ phases travel through them, so a 26-qubit, 12-layer circuit
    # needs few passes
    from tyxonq_amd import serving as S

    req = S.Request(0, 26, 12, "tfim", 1.0, np.random.default_rng(1).uniform(0, 6.28, 624))
    info = nat.inspect(26, S.request_ops(req))
    assert info["passes"] <= 34, info  # 46 with the CX as a general 2x2 step
    small = S.Request(0, 12, 4, "tfim", 1.0, np.random.default_rng(2).uniform(0, 6.28, 96))
    v = rand_state(np.random.default_rng(3), 12)
    ops = S.request_ops(small)
    want = v.copy()
    for row in ops:
        m = np.array(row["m"][:8:2] + 1j * row["m"][1:8:2]).reshape(2, 2)
        item = ("m1", 11 - int(row["t"][0]), m, int(row["cmask"]), int(row["cval"]))
        want = dense_apply_items(12, [item], want)
    np.testing.assert_allclose(nat.emulate(12, ops, v.astype(np.complex64)), want, atol=5e-5)


@pytest.mark.parametrize("n", [1, 2, 5, 8, 10])
def test_small_states_run_padded(n):
    # fewer than 11 qubits: planned as an 11-qubit state whose extra high bits
    # stay zero (the emulator checks that they do) and copied back
    rng = np.random.default_rng(900 + n)
    ops = random_circuit_ops(rng, n, 120, p_two=0.4 if n > 1 else 0.0)
    v = rand_state(rng, n)
    want, _ = O.simulate(n, ops, mode="run", init=v, reference_dispatch=False)
    np.testing.assert_allclose(_emulated(n, ops, v), want, atol=2e-5)
    if n >= 2:
        info = nat.inspect(n, lower(tq.Circuit(n, ops=list(ops)), mode="run",
                                    reference_dispatch=False)[0].segments[0][1])
        assert info["n"] == n and info["passes"] <= 4


def _commuting_group(rng, n, x_qubits, size):
    """`size` Pauli strings on x_qubits with X/Y letters whose Y counts all
    have one parity (so they commute pairwise) and a shared Z string outside."""
    zq = [q for q in range(n) if q not in x_qubits][: int(rng.integers(0, 4))]
    par = int(rng.integers(2))
    out, seen = [], set()
    while len(out) < size:
        letters = tuple("XY"[int(rng.integers(2))] for _ in x_qubits)
        if letters.count("Y") % 2 != par or letters in seen:
            continue
        seen.add(letters)
        out.append(tuple(zip(letters, x_qubits)) + tuple(("Z", q) for q in zq))
    return out


@pytest.mark.parametrize("width,size", [(2, 2), (4, 8), (3, 4), (5, 16)])
def test_grouped_commuting_rotations_emulated(width, size):
    """Runs of commuting rotations with one flip mask become one grouped step
    (a two-level rotation per pair class); compared with the oracle's
    rotation-by-rotation product, interleaved with gates that break runs."""
    n = 13
    rng = np.random.default_rng(10 * width + size)
    ops = []
    for _ in range(6):
        xq = [int(q) for q in rng.choice(n, width, replace=False)]
        for fac in _commuting_group(rng, n, xq, size):
            ops.append(("pauli_rot", fac, float(rng.normal(0, 0.7))))
        ops.append(("ry", int(rng.integers(n)), 0.4))
    v = rand_state(rng, n)
    want, _ = O.simulate(n, ops, mode="run", init=v, reference_dispatch=False)
    np.testing.assert_allclose(_emulated(n, ops, v), want, atol=2e-5)
    info = nat.inspect(n, lower(tq.Circuit(n, ops=ops), mode="run",
                                reference_dispatch=False)[0].segments[0][1])
    assert info["gates_after_merge"] <= 6 * 2 + 1  # each run became one gate


def test_ucc_circuit_emulated():
    from tyxonq_amd import chem

    ns, no = 5, 2
    n = 2 * ns
    ex, ids = chem.uccsd_excitations(no, ns - no)
    th = np.random.default_rng(3).normal(0, 0.3, size=max(ids) + 1)
    ans = chem.UCCAnsatz(n, (no, no), ex, ids)
    c = chem.ucc_circuit(list(th), n, (no, no), ex, ids)
    want, _ = O.simulate(n, c.ops)
    v = np.zeros(1 << n, dtype=np.complex64)
    v[0] = 1
    got = nat.emulate(n, ans.ops(th), v)
    np.testing.assert_allclose(got, want, atol=2e-6)
    assert nat.inspect(n, ans.ops(th))["gates_after_merge"] <= len(ex) + 2 * no
