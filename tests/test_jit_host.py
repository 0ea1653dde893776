"""Generated pass kernels compile with the installation's run-time compiler
(hiprtc), on the host, with no GPU.

tqsv_plan_precompile plans a circuit on the host and compiles every distinct
pass kernel of the plan (and of its reverse sweep): a generated source that
hiprtc rejects fails here instead of on first use on a GPU.  The circuits
cover every step kind the generator emits (rotations, phases, sign steps,
controlled swaps, general and controlled 1-qubit matrices, dense 2-qubit
matrices, Pauli rotations, rotation groups and their sparse phases, the
cross and group reductions of dual passes) and both index widths (32-bit up
to 32 qubits, 64-bit above).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import tyxonq_amd as tq
from tests.gpu_util import ops_from_items, random_circuit_ops, random_matrix_items
from tyxonq_amd import _native as nat
from tyxonq_amd.engine import lower


def _have_hiprtc() -> bool:
    root = os.environ.get("ROCM_PATH") or "/opt/rocm"
    return os.path.exists(os.path.join(root, "lib", "libhiprtc.so.7"))


pytestmark = pytest.mark.skipif(not _have_hiprtc(), reason="no hiprtc in this installation")


@pytest.fixture(autouse=True)
def _private_compiler_cache(tmp_path, monkeypatch):
    # the compiler's code-object cache of these tests stays in a temporary folder
    monkeypatch.setenv("AMD_COMGR_CACHE_DIR", str(tmp_path / "comgr"))


def _ops(n, ops, track_grad=False):
    L, _, _ = lower(tq.Circuit(n, ops=ops), mode="run", reference_dispatch=False,
                    track_grad=track_grad)
    return L.segments[0][1]


@pytest.mark.parametrize("env", [{}, {"TQSV_PIPE": "0"}, {"TQSV_DUAL1": "0"}])
def test_random_gate_circuit_forward_and_reverse(monkeypatch, env):
    # TQSV_PIPE=0: no register-staged tiles in gate passes; TQSV_DUAL1=0: dual
    # passes with two tile buffers
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    n = 15
    ops = random_circuit_ops(np.random.default_rng(5), n, 70)
    k = nat.precompile(n, _ops(n, ops), adjoint=True)
    assert k >= 2


def test_matrix_items_with_controls():
    n = 14
    items = random_matrix_items(np.random.default_rng(8), n, 40)
    assert nat.precompile(n, ops_from_items(n, items)) >= 1


def test_parametrised_ansatz_adjoint_cross_reductions():
    import torch

    from tyxonq_amd import bench_workloads as W

    n, layers = 16, 3
    tt = torch.zeros(2 * n * layers, dtype=torch.float64, requires_grad=True)
    circ, _ = W.ryrz_ring_ansatz(n, layers, params=[tt[i] for i in range(tt.numel())])
    L, _, _ = lower(circ, track_grad=True)
    assert nat.precompile(n, L.segments[0][1], adjoint=True) >= 2


def test_pauli_rotations_and_ucc_groups_adjoint():
    from tyxonq_amd import chem

    ns, no = 7, 3
    n = 2 * ns
    ex, ids = chem.uccsd_excitations(no, ns - no)
    ans = chem.UCCAnsatz(n, (no, no), ex, ids)
    th = np.random.default_rng(2).normal(0, 0.1, size=ans.nparams)
    assert nat.precompile(n, ans.ops(th), adjoint=True, max_cost=100000) >= 2
    # single Pauli rotations with distinct flip masks (no groups)
    rng = np.random.default_rng(4)
    rots = [("pauli_rot", tuple(("XYZ"[int(rng.integers(3))], int(q))
                                for q in rng.choice(n, 3, replace=False)), 0.3 + 0.1 * j)
            for j in range(20)]
    assert nat.precompile(n, _ops(n, rots)) >= 1


def test_wide_state_uses_64_bit_indices():
    n = 33
    ops = [("rx", q, 0.1 * q) for q in range(n)] + [("cz", q, q + 1) for q in range(0, n - 1, 2)]
    assert nat.precompile(n, _ops(n, ops)) >= 1


def test_expectation_passes_of_a_wide_state():
    """The 500-term 2-local Pauli sum on 30 qubits: read-only passes with two
    tile buffers filled by asynchronous copies (states of 2^25+ amplitudes),
    and the same sum on 33 qubits (64-bit indices)."""
    from tyxonq_amd import bench_workloads as W
    from tyxonq_amd.pauli import PauliSum

    for n in (30, 33):
        x, z, c = PauliSum(n, W.random_2local_terms(n, 120)).to_masks()
        assert nat.ham_precompile(n, x, z, c) >= 1


def test_compiles_after_the_environment_grew(monkeypatch):
    """The compiler runs in a linker namespace with a C library of its own.
    Environment variables added after it was loaded move the process's
    environment array (and free the old one); the next compile must still
    find a valid environment there (it once read the freed array and ended
    in a segmentation fault)."""
    n = 12
    assert nat.precompile(n, _ops(n, random_circuit_ops(np.random.default_rng(1), n, 30))) >= 1
    for rnd in range(3):
        for k in range(150):
            monkeypatch.setenv(f"TQSV_TEST_ENV_{rnd}_{k}", "x" * 40)
        junk = [bytearray(64 * (j % 50 + 1)) for j in range(20000)]  # reuse freed blocks
        m = 13 + rnd
        ops = _ops(m, random_circuit_ops(np.random.default_rng(10 + rnd), m, 30))
        assert nat.precompile(m, ops) >= 1
        del junk
        for k in range(0, 150, 2):
            monkeypatch.delenv(f"TQSV_TEST_ENV_{rnd}_{k}")
