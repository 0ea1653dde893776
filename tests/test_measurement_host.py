"""Measurement feed on the host: grouping, the measurement stage, shot plans,
parameter-shift pairs, counts parities and the basis-rotation rows."""
from __future__ import annotations

import math

import numpy as np
import pytest

import tyxonq_amd as tq
from oracle import statevector as O
from oracle import dense as D
from tyxonq_amd import _native as nat
from tyxonq_amd import measurement as M
from tyxonq_amd.compiler import build_plan, compile_native
from tyxonq_amd.postprocessing import term_expectation_from_counts


def test_measurement_stage_groups_non_overlapping_expectations():
    # tests_core_module/test_measurement_rewrite_grouping.py:6-29
    circ = tq.Circuit(2, ops=[])
    ms = [M.Expectation("Z", (0,)), M.Expectation("X", (1,)), M.Expectation("ZZ", (0, 1))]
    out = build_plan(["rewrite/measurement"]).execute_plan(circ, device_rule={}, measurements=ms)
    assert out is circ
    groups = circ.metadata["measurement_groups"]
    assert len(groups) == 2 and sum(len(g["items"]) for g in groups) == 3
    assert {g["basis"] for g in groups} == {"pauli"}
    g0 = next(g for g in groups if len(g["items"]) == 2)
    assert set(g0["basis_map"]) == {0, 1}
    assert all("estimated_settings" in g and "estimated_shots_per_group" in g for g in groups)


def test_group_based_shot_scheduling():
    # tests_core_module/test_shot_scheduler_integration.py:7-19
    circ = tq.Circuit(2, ops=[])
    ms = [M.Expectation("Z", (0,)), M.Expectation("X", (1,)), M.Expectation("ZZ", (0, 1))]
    circ = build_plan(["rewrite/measurement"]).execute_plan(circ, device_rule={}, measurements=ms)
    plan = M.schedule(circ, total_shots=1000)
    assert len(plan["segments"]) >= 2
    assert sum(s["shots"] for s in plan["segments"]) == 1000
    assert all("basis" in s and "wires" in s for s in plan["segments"])
    # weights 2*2 and 1*2 -> 667 + 333; max_shots_per_job splits, batch ids tag
    split = M.schedule(circ, total_shots=1000,
                       device_rule={"max_shots_per_job": 300, "supports_batch": True,
                                    "max_segments_per_batch": 2})
    assert [s["shots"] for s in split["segments"]] == [300, 300, 67, 300, 33]
    assert [s["batch_id"] for s in split["segments"]] == [0, 0, 1, 1, 2]
    assert M.schedule(circ, [10, 20])["segments"] == [{"shots": 10}, {"shots": 20}]


def test_pipeline_stages_and_errors():
    # tests_core_module/test_compiler_pipeline.py:7-39
    pl = build_plan(["decompose", "rewrite/measurement", "layout", "scheduling",
                     "scheduling/shot_scheduler"])
    circ = tq.Circuit(2, ops=[("h", 0)])
    assert pl.execute_plan(circ, device_rule={}) is circ
    with pytest.raises(ValueError):
        build_plan(["unknown_stage"])
    with pytest.raises(ValueError):
        build_plan(["scheduling/shot_scheduler"]).execute_plan(circ, shot_plan=[0, 3])
    c1 = tq.Circuit(1, ops=[("rz", 0, 0.1), ("measure_z", 0)])
    out = build_plan(["gradients/parameter_shift"]).execute_plan(c1, device_rule={}, grad_op="rz")
    g = out.metadata["gradients"]["rz"]
    assert g["plus"].ops[0][2] == pytest.approx(0.1 + math.pi / 2)
    assert g["minus"].ops[0][2] == pytest.approx(0.1 - math.pi / 2)
    assert g["meta"]["coeff"] == 0.5


def test_compile_with_plan_and_shot_schedule():
    c = tq.Circuit(2).h(0).cx(0, 1)
    ham = [(0.5, [("Z", 0)]), (0.3, [("X", 0), ("X", 1)]), (1.0, [])]
    with pytest.warns(UserWarning):
        low, meta = compile_native(c, {"compile_plan": ["rewrite/measurement"],
                                       "hamiltonian_terms": ham, "total_shots": 900})
    groups = low.metadata["measurement_groups"]
    assert low.metadata["measurement_context"]["identity_const"] == 1.0
    assert [g.get("source") for g in groups] == [None, "hamiltonian", "hamiltonian"]
    assert meta["compile_plan"][-1] == "rewrite/measurement"
    assert sum(s["shots"] for s in meta["job_plan"]["segments"]) == 900


def test_hamiltonian_grouping_known_answer():
    # hamiltonian_grouping.py:43-55: 0.5 Z0 + 0.3 X0 X1 + 1.0 I
    qop = {((0, "Z"),): 0.5, ((0, "X"), (1, "X")): 0.3, (): 1.0}
    const, groups = M.group_qubit_operator_terms(qop, 2)
    assert const == 1.0 and len(groups) == 2
    assert groups[("Z", "I")] == [(((0, "Z"),), 0.5)]
    assert groups[("X", "X")] == [(((0, "X"), (1, "X")), 0.3)]
    with pytest.raises(ValueError):
        M.group_qubit_operator_terms({((0, "Z"),): 0.5 + 1e-6j}, 1)
    const2, groups2 = M.group_hamiltonian_pauli_terms(
        [(0.5, [("Z", 0)]), (0.3, [("X", 0), ("X", 1)]), (1.0, []), (0.2, [("Z", 0)])], 2)
    assert const2 == 1.0 and groups2[("Z", "I")] == [(((0, "Z"),), 0.5), (((0, "Z"),), 0.2)]


def test_z_products_match_the_bitstring_rule():
    rng = np.random.default_rng(4)
    n = 7
    idx = np.unique(rng.integers(0, 1 << n, 60)).astype(np.uint64)
    cnt = rng.integers(1, 50, idx.size)
    counts = {format(int(i), f"0{n}b"): int(c) for i, c in zip(idx, cnt)}
    terms = [((0, "Z"),), ((1, "X"), (4, "Y")), ((0, "Z"), (3, "Z"), (6, "X")), ()]
    got = M.z_products(idx, cnt, terms, n)
    want = [term_expectation_from_counts(counts, [q for q, _ in t]) for t in terms]
    np.testing.assert_allclose(got, want, atol=1e-14)


@pytest.mark.parametrize("bases", [("X", "I", "Y"), ("Y", "Y", "Z"), ("I", "I", "I")])
def test_basis_rotation_rows(bases):
    n = 3
    rng = np.random.default_rng(11)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    got = nat.emulate(n, M.basis_rotation_rows(bases, n), psi.astype(np.complex64))
    want = psi.astype(np.complex128)
    for q, P in enumerate(bases):
        if P == "X":
            want = O.apply_1q(want, O.gate_h(), q, n)
        elif P == "Y":
            want = O.apply_1q(O.apply_1q(want, O.gate_sd(), q, n), O.gate_h(), q, n)
    np.testing.assert_allclose(got, want, atol=2e-6)
    # measuring in the rotated basis gives the Pauli expectation of the original state
    term = [(P, q) for q, P in enumerate(bases) if P != "I"]
    if term:
        p = np.abs(want) ** 2
        idx = np.arange(1 << n, dtype=np.uint64)
        zp = M.z_products(idx, p * 1e6, [tuple((q, P) for P, q in term)], n)[0]
        exact = np.real(np.vdot(psi, D.pauli_sum_dense([(1.0, term)], n) @ psi))
        assert zp == pytest.approx(exact, abs=1e-9)


def test_bitstrings_are_big_endian_binary():
    from tyxonq_amd.engine import bitstrings

    for n in (1, 7, 30, 63):
        idx = np.random.default_rng(n).integers(0, 2 ** min(n, 62), 40).astype(np.uint64)
        assert bitstrings(idx, n) == [format(int(i), f"0{n}b") for i in idx]
    assert bitstrings(np.zeros(0, dtype=np.uint64), 4) == []


def test_shared_prefix_keys():
    import torch

    from tyxonq_amd.engine import _common_prefix, _op_key, _op_keys

    a = tq.Circuit(3).h(0).rx(1, 0.5).cx(0, 1)
    b = tq.Circuit(3).h(0).rx(1, 0.5).cz(1, 2)
    assert _common_prefix(_op_keys(a), _op_keys(b)) == 2
    # tensor angles compare by value; a different angle ends the prefix
    t1 = tq.Circuit(1, ops=[("rz", 0, torch.tensor(0.25)), ("h", 0)])
    t2 = tq.Circuit(1, ops=[("rz", 0, torch.tensor(0.25)), ("x", 0)])
    t3 = tq.Circuit(1, ops=[("rz", 0, torch.tensor(0.3)), ("h", 0)])
    assert _common_prefix(_op_keys(t1), _op_keys(t2)) == 1
    assert _common_prefix(_op_keys(t1), _op_keys(t3)) == 0
    # stochastic and measurement-type ops are never shared
    for op in (("kraus", 0, "k"), ("reset", 0), ("project_z", 0, 1)):
        assert _op_key(op) is None
    k = [("h", 0), ("kraus", 0, "k"), ("h", 0)]
    assert _common_prefix([_op_key(o) for o in k], [_op_key(o) for o in k]) == 1
    # pauli rotations with factor lists and array angles
    p = ("pauli_rot", [("X", 0), ("Z", 1)], np.float64(0.7))
    assert _op_key(p) == _op_key(("pauli_rot", [("X", 0), ("Z", 1)], np.float64(0.7)))
