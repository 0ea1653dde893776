"""Full-size parity against complex128 goldens (GPU).

tests/golden/fullsize_c128.json holds complex128 results of the oracle's C
restatement at the benchmark sizes (written by
tests/golden/make_fullsize_goldens.py; the reference cannot run them).  Each
test runs the HIP path exactly as bench.py configures it and checks:

  * 30-qubit brickwork (20 layers, generated streaming kernels, pass budget
    260): |dE| <= 1e-6, every one of the 500 <P_k> within 2e-6, the norm
    within 1e-6; the generated and the interpreting kernels give the same
    state (||dpsi|| <= 1e-5); the inverse circuit returns |0...0>.
  * 22-qubit RY.RZ + CX-ring adjoint (528 parameters): |dE| <= 1e-6 and
    max|dgrad| <= 1e-5 over all 528 components.
  * 20-spin-orbital UCCSD (450 parameters, 6700 Pauli rotations, 14251-term
    Jordan-Wigner Hamiltonian): |dE| <= 1e-6 and max|dgrad| <= 1e-5.
  * 26-qubit RY.RZ + CX-ring adjoint (2 layers, several tiles per dual-pass
    work-group): max|dgrad| <= 1e-5 and |dE| <= 4e-7 |E| (|E| = 10.7).
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "fullsize_c128.json").read_text())


def _need(name):
    if name not in GOLD:
        pytest.skip(f"golden {name} not generated yet")
    return GOLD[name]


@pytest.fixture(scope="module")
def brickwork30(gpu):
    import torch

    from tyxonq_amd import _native as nat
    from tyxonq_amd import bench_workloads as W
    from tyxonq_amd.engine import lower
    from tyxonq_amd.pauli import PauliSum

    g = _need("brickwork30")
    n, layers = g["qubits"], g["layers"]
    circ = W.brickwork_circuit(n, layers, g["seed"])
    terms = W.random_2local_terms(n, g["pauli_terms"], g["seed"])
    L, _, _ = lower(circ)
    ops = L.segments[0][1]
    plan = nat.Plan(n, ops, max_cost=260)
    x, z, c = PauliSum(n, terms).to_masks()
    ham = nat.Ham(n, x, z, c)
    psi = torch.empty(1 << n, dtype=torch.complex64, device=nat.device())
    nat.init_basis(psi, 0)
    plan.run(psi)
    yield {"n": n, "circ": circ, "ops": ops, "plan": plan, "ham": ham, "psi": psi, "gold": g}
    del psi
    torch.cuda.empty_cache()


def test_brickwork30_energy_terms_norm(brickwork30):
    b = brickwork30
    e, nrm, per_term = b["ham"].expval(b["psi"], per_term=True)
    g = b["gold"]
    print(f"30q |dE| = {abs(e - g['energy']):.3e}, |dnorm| = {abs(nrm - g['norm2']):.3e}, "
          f"max |d<P>| = {np.max(np.abs(per_term - np.asarray(g['term_values']))):.3e}")
    assert abs(nrm - g["norm2"]) <= 1e-6
    np.testing.assert_allclose(per_term, g["term_values"], atol=2e-6)
    assert abs(e - g["energy"]) <= 1e-6


def test_brickwork30_generated_equals_interpreting(brickwork30):
    import torch

    from tyxonq_amd import _native as nat

    b = brickwork30
    assert b["plan"].info()["passes"] > 0
    ref = torch.empty_like(b["psi"])
    nat.init_basis(ref, 0)
    interp = nat.Plan(b["n"], b["ops"], max_cost=260, jit=-1)
    interp.run(ref)
    diff = float(torch.linalg.vector_norm(ref - b["psi"]).item())
    print(f"30q ||psi_generated - psi_interpreting|| = {diff:.3e}")
    assert diff <= 1e-5
    del ref


def test_brickwork30_inverse_returns_basis_state(brickwork30):
    import torch

    from tyxonq_amd import _native as nat
    from tyxonq_amd.circuit import Circuit
    from tyxonq_amd.engine import lower

    b = brickwork30
    inv = Circuit(b["n"])
    for op in reversed(b["circ"].ops):
        if op[0] in ("rx", "rz"):
            inv.ops.append((op[0], op[1], -float(op[2])))
        else:
            inv.ops.append(op)
    L, _, _ = lower(inv)
    work = b["psi"].clone()
    nat.Plan(b["n"], L.segments[0][1], max_cost=260).run(work)
    p0 = float(work[0].abs().item() ** 2)
    rest = float(torch.linalg.vector_norm(work[1:]).item())
    print(f"30q U^dagger U |0>: |amp0|^2 = {p0:.9f}, ||rest|| = {rest:.3e}")
    assert abs(p0 - 1.0) <= 2e-6 and rest <= 2e-3
    del work


def test_adjoint22_energy_and_full_gradient(gpu):
    import torch

    from tyxonq_amd import _native as nat
    from tyxonq_amd import bench_workloads as W
    from tyxonq_amd.engine import lower
    from tyxonq_amd.pauli import PauliSum

    g = _need("adjoint22")
    n, layers = g["qubits"], g["layers"]
    tt = torch.tensor(np.random.default_rng(g["seed"]).normal(0.0, 0.5, size=2 * n * layers),
                      dtype=torch.float64, requires_grad=True)
    circ, _ = W.ryrz_ring_ansatz(n, layers, params=[tt[i] for i in range(tt.numel())])
    L, _, _ = lower(circ, track_grad=True)
    plan = nat.Plan(n, L.segments[0][1])
    x, z, c = PauliSum(n, W.heisenberg_ring_terms(n, 0.5)).to_masks()
    ham = nat.Ham(n, x, z, c)
    psi = torch.empty(1 << n, dtype=torch.complex64, device=nat.device())
    lam = torch.empty_like(psi)
    nat.init_basis(psi, 0)
    e, grad = plan.adjoint(ham, psi, lam)
    gref = np.asarray(g["grad"])
    print(f"22q adjoint |dE| = {abs(e - g['energy']):.3e}, "
          f"max|dgrad| = {np.max(np.abs(grad - gref)):.3e} (max|g| {np.max(np.abs(gref)):.3f})")
    np.testing.assert_allclose(gref[g["ps_indices"]], g["ps_grad"], atol=1e-10)
    assert abs(e - g["energy"]) <= 1e-6
    np.testing.assert_allclose(grad, gref, atol=1e-5)


@pytest.mark.parametrize("dual1", ["1", "0"])
def test_adjoint26_dual_passes_loop_over_tiles(gpu, monkeypatch, dual1):
    """26 qubits, 2 layers, 104 parameters: 8192 tiles for at most 2048
    dual-pass work-groups, so each work-group of a generated dual pass runs
    four tiles in turn (one tile buffer, or two with TQSV_DUAL1=0).  The
    generated sweep and the interpreting kernel's sweep of the same plan are
    each held to the complex128 result: max|dgrad| <= 1e-5, and the energy
    relative to its size, |dE| <= 4e-7 |E|.  This Hamiltonian's energy is
    10.7, and complex64 amplitudes carry about 1e-7 relative error, so the
    1e-6 absolute bound the benchmark configurations meet (|E| <= 1 there)
    is out of reach here: measured 2.4e-6 (generated) and 3.4e-6
    (interpreting); the norm drifts by 3.7e-7."""
    import torch

    from tyxonq_amd import _native as nat
    from tyxonq_amd import bench_workloads as W
    from tyxonq_amd.engine import lower
    from tyxonq_amd.pauli import PauliSum

    g = _need("adjoint26")
    monkeypatch.setenv("TQSV_DUAL1", dual1)
    n, layers = g["qubits"], g["layers"]
    th = np.random.default_rng(g["seed"]).normal(0.0, 0.5, size=2 * n * layers)
    tt = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    circ, _ = W.ryrz_ring_ansatz(n, layers, params=[tt[i] for i in range(tt.numel())])
    L, _, _ = lower(circ, track_grad=True)
    ops = L.segments[0][1]
    x, z, c = PauliSum(n, W.heisenberg_ring_terms(n, 0.5)).to_masks()
    ham = nat.Ham(n, x, z, c)
    psi = torch.empty(1 << n, dtype=torch.complex64, device=nat.device())
    lam = torch.empty_like(psi)
    gref = np.asarray(g["grad"])
    for jit in (1, -1):
        plan = nat.Plan(n, ops, jit=jit)
        nat.init_basis(psi, 0)
        e, grad = plan.adjoint(ham, psi, lam)
        assert plan.info()["reverse_generated"] == (1 if jit == 1 else 0)
        del plan
        print(f"26q adjoint (jit {jit}, TQSV_DUAL1={dual1}) |dE| = {abs(e - g['energy']):.3e}, "
              f"max|dgrad| = {np.max(np.abs(grad - gref)):.3e} (E {g['energy']:.6f}, "
              f"max|g| {np.max(np.abs(gref)):.3f})")
        assert abs(e - g["energy"]) <= 4e-7 * abs(g["energy"])
        np.testing.assert_allclose(grad, gref, atol=1e-5)
    del psi, lam


@pytest.mark.parametrize("jit", [0, 1])
def test_ucc20_energy_and_full_gradient(gpu, jit):
    """jit 0: the automatic choice (interpreting kernels at 20 qubits); jit 1:
    generated kernels, with the rotation groups' sparse phases."""
    import torch

    from tyxonq_amd import _native as nat
    from tyxonq_amd import chem
    from tyxonq_amd.pauli import PauliSum

    g = _need("ucc20")
    ns = g["spin_orbitals"] // 2
    no = ns // 2
    n = 2 * ns
    h, gi = chem.synthetic_integrals(ns, g["seed"])
    terms = chem.molecular_hamiltonian(h, gi)
    ex, ids = chem.uccsd_excitations(no, ns - no)
    ans = chem.UCCAnsatz(n, (no, no), ex, ids)
    th = np.random.default_rng(g["seed"]).normal(0.0, g["theta_sigma"], size=ans.nparams)
    x, z, c = PauliSum(n, terms).to_masks()
    ham = nat.Ham(n, x, z, c)
    plan = nat.Plan(n, ans.ops(th), max_cost=100000, jit=jit)
    psi = torch.empty(1 << n, dtype=torch.complex64, device=nat.device())
    lam = torch.empty_like(psi)
    nat.init_basis(psi, 0)
    e, grad = plan.adjoint(ham, psi, lam)
    info = plan.info()
    assert (info["generated"], info["reverse_generated"]) == (jit, jit)
    gref = np.asarray(g["grad"])
    print(f"ucc20 jit={jit} |dE| = {abs(e - g['energy']):.3e}, "
          f"max|dgrad| = {np.max(np.abs(grad - gref)):.3e} (max|g| {np.max(np.abs(gref)):.3f})")
    np.testing.assert_allclose(gref[g["fd_indices"]], g["fd_grad"], atol=1e-7)
    assert abs(e - g["energy"]) <= 1e-6
    np.testing.assert_allclose(grad, gref, atol=1e-5)
