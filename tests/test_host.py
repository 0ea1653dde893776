"""Host-side logic of the engine (CPU only): the Circuit IR surface, lowering
of op tuples to tqsv_op records, Pauli masks, the compile passes and the
post-processing router."""
from __future__ import annotations

import math
import warnings

import numpy as np
import pytest

import tyxonq_amd as tq
from tyxonq_amd import _native as nat
from tyxonq_amd import gates as G
from tyxonq_amd.compiler import compile_native
from tyxonq_amd.engine import lower
from tyxonq_amd.pauli import PauliSum
from tyxonq_amd.postprocessing import apply_postprocessing
from oracle import statevector as O


def test_circuit_builders_and_validation():
    c = tq.Circuit(3)
    assert c.h(0).cx(0, 1).rz(2, 0.3).RX(1, 0.1).cnot(1, 2) is c
    assert c.ops == [("h", 0), ("cx", 0, 1), ("rz", 2, 0.3), ("rx", 1, 0.1), ("cx", 1, 2)]
    with pytest.raises(ValueError, match="out of range"):
        tq.Circuit(2, ops=[("h", 5)])
    with pytest.raises(ValueError, match="Matrix shape.*incompatible"):
        tq.Circuit(2).unitary(0, matrix=np.eye(4))
    with pytest.raises(ValueError, match="Matrix shape.*incompatible"):
        tq.Circuit(2).unitary(0, 1, matrix=np.eye(2))
    with pytest.raises(ValueError, match="Invalid qubit index"):
        tq.Circuit(2).unitary(5, matrix=np.eye(2))
    with pytest.warns(UserWarning):
        tq.Circuit(1).reset(0)


def test_json_roundtrip():
    c = tq.Circuit(2).h(0).rx(1, 0.25).cx(0, 1)
    c2 = tq.Circuit.from_json_str(c.to_json_str())
    assert c2.num_qubits == 2 and c2.ops == c.ops


def test_gate_matrices_match_oracle():
    t = 0.77
    pairs = [(G.gate_h(), O.gate_h()), (G.gate_rx(t), O.gate_rx(t)), (G.gate_ry(t), O.gate_ry(t)),
             (G.gate_rz(t), O.gate_rz(t)), (G.gate_s(), O.gate_s()), (G.gate_sd(), O.gate_sd()),
             (G.gate_cx_4x4(), O.gate_cx()), (G.gate_cz_4x4(), O.gate_cz()),
             (G.gate_iswap_4x4(), O.gate_iswap()), (G.gate_swap_4x4(), O.gate_swap()),
             (G.gate_rxx(t), O.gate_rxx(t)), (G.gate_ryy(t), O.gate_ryy(t)),
             (G.gate_rzz(t), O.gate_rzz(t)), (G.gate_cry_4x4(t), O.gate_cry(t))]
    for a, b in pairs:
        np.testing.assert_allclose(a, b, atol=1e-15)
    np.testing.assert_allclose(G.build_controlled_unitary(G.gate_x(), 1), O.gate_cx(), atol=0)
    np.testing.assert_allclose(G.build_controlled_unitary(G.gate_x(), 2, [1, 0]),
                               O.build_controlled_unitary(O.X, 2, [1, 0]), atol=0)


def test_lowering_bits_params_and_skips():
    import torch

    th = torch.tensor(0.4, dtype=torch.float64, requires_grad=True)
    c = tq.Circuit(3).h(0).ry(1, th).cx(0, 2).z(1).rzz(0, 2, 0.2).measure_z(0)
    L, meas, _ = lower(c, mode="state", track_grad=True)
    ops = L.segments[0][1]
    assert list(ops["kind"]) == [nat.OP_MAT1, nat.OP_MAT1, nat.OP_MAT1, nat.OP_MAT2]
    assert ops["t"][0][0] == 2  # qubit 0 -> bit n-1-0
    assert ops["t"][1][0] == 1
    assert ops["param"][1] == 0 and ops["gx"][1] == 2 and ops["gz"][1] == 2  # Y generator on bit 1
    assert ops["cmask"][2] == 4 and ops["cval"][2] == 4 and ops["t"][2][0] == 0  # cx: X on bit 0
    assert tuple(ops["t"][3]) == (2, 0) and ops["gz"][3] == 5
    assert L.params == [th]
    assert meas == [0]
    # z is skipped by the reference dispatch, applied when asked
    L2, _, _ = lower(c, mode="state", reference_dispatch=False)
    assert len(L2.segments[0][1]) == 5


def test_lowering_segments_on_measurement_ops():
    c = tq.Circuit(2).h(0).mid_measurement(0, 1).h(1)
    L, _, _ = lower(c)
    assert [s[0] for s in L.segments] == ["ops", "project", "ops"]
    assert L.segments[1][1:] == (1, 1)


def test_pauli_masks_and_formats():
    n = 3
    ps = PauliSum(n, [(0.5, [("X", 0), ("Z", 2)]), (-1.0, [("Y", 1)]), (2.0, [])])
    x, z, c = ps.to_masks()
    assert list(x) == [4, 2, 0] and list(z) == [1, 2, 0] and list(c) == [0.5, -1.0, 2.0]
    of = {((0, "X"), (2, "Z")): 0.5, ((1, "Y"),): -1.0, (): 2.0}
    assert PauliSum.coerce(of, n).to_masks()[0].tolist() == [4, 2, 0]
    codes = ([[1, 0, 3], [0, 2, 0]], [0.5, -1.0])
    assert PauliSum.coerce(codes, n).to_masks()[1].tolist() == [1, 2]
    with pytest.raises(ValueError, match="not real"):
        PauliSum(1, [(1j, [("X", 0)])])
    with pytest.raises(ValueError, match="twice"):
        PauliSum(2, [(1.0, [("X", 0), ("Z", 0)])])
    H = ps.to_dense()
    assert H.shape == (8, 8) and np.allclose(H, H.conj().T)


def test_compile_native_passes():
    c = tq.Circuit(2).x(0).y(1)
    c.unitary(0, matrix=np.eye(2))
    with pytest.warns(UserWarning, match="auto-added"):
        low, meta = compile_native(c, {})
    names = [op[0] for op in low.ops]
    assert names[:2] == ["rx", "y"]
    assert low.ops[0][2] == pytest.approx(math.pi)
    assert names.count("measure_z") == 2
    assert not hasattr(low, "_unitary_cache")  # dropped by the rebuild, as in the reference
    assert meta["compile_plan"] == ["rewrite/auto_measure", "rewrite/gates_transform"]
    assert meta["job_plan"] is None


def test_postprocessing_router():
    counts = {"00": 400, "01": 100, "10": 300, "11": 200}
    assert apply_postprocessing({"result": counts}, {"method": None}) == {"method": None, "result": None}
    r = apply_postprocessing({"result": counts}, {"method": "expval_pauli_term", "idxs": [0]})
    assert r["result"] == pytest.approx(0.0)  # (400+100-300-200)/1000
    items = [(((0, "Z"),), 0.5), (((0, "Z"), (1, "Z")), 1.0)]
    r = apply_postprocessing({"result": counts}, {"method": "expval_pauli_sum", "items": items})
    zz = (400 - 100 - 300 + 200) / 1000
    assert r["result"]["energy"] == pytest.approx(0.5 * 0.0 + zz)
    probs = np.array([0.4, 0.1, 0.3, 0.2])
    r = apply_postprocessing({"result": {}, "expectations": {"Z0": 0.0}, "probabilities": probs,
                              "metadata": {"num_qubits": 2}},
                             {"method": "expval_pauli_sum", "items": items})
    assert r["result"]["energy"] == pytest.approx(zz)


def test_backend_selection():
    prev = tq.get_backend().name
    try:
        assert tq.set_backend("torch").name == "pytorch"
        with tq.use_backend("numpy") as b:
            assert b.name == "numpy" and tq.get_backend().name == "numpy"
        assert tq.get_backend().name == "pytorch"
        with pytest.raises(ValueError):
            tq.set_backend("jax")
    finally:
        tq.set_backend(prev)


def test_value_and_grad_cpu_function():
    import torch

    vg = tq.value_and_grad(lambda p: torch.sum(torch.sin(p) ** 2))
    v, g = vg(np.array([0.3, 1.1]))
    assert v == pytest.approx(math.sin(0.3) ** 2 + math.sin(1.1) ** 2)
    np.testing.assert_allclose(g, [math.sin(0.6), math.sin(2.2)], atol=1e-12)


def test_parameter_shift_gradient_convention():
    from tyxonq_amd.kernels import parameter_shift_gradient

    th = np.array([0.3, -1.1, 2.0])
    # E = cos t0 * cos t1 + sin t2 (each term generated by a +-1/2 Pauli rotation)
    f = lambda p: np.cos(p[0]) * np.cos(p[1]) + np.sin(p[2])
    g = parameter_shift_gradient(f, th)
    want = [-np.sin(th[0]) * np.cos(th[1]), -np.cos(th[0]) * np.sin(th[1]), np.cos(th[2])]
    np.testing.assert_allclose(g, want, atol=1e-12)
    with pytest.raises(ValueError):
        parameter_shift_gradient(f, np.zeros((2, 2)))


def test_lowering_attenuates_only_wires():
    """engine.py:56-141: depolarizing noise attenuates <Z> of the op's wires only."""
    from tyxonq_amd.circuit import Circuit

    noise = {"type": "depolarizing", "p": 0.3}
    _, _, z = lower(Circuit(2).rx(0, 1), mode="run", noise=noise)
    assert z == pytest.approx([0.6, 1.0])
    _, _, z = lower(Circuit(3).rz(2, 5).cx(0, 1).rzz(1, 2, 3), mode="run", noise=noise)
    assert z == pytest.approx([0.6, 0.36, 0.36])
