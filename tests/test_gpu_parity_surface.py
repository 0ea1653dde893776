"""Surface parity on the GPU engine (GPU): the shots=0 driver payload, the
depolarizing attenuation wires, perfect_sampling, the plain per-parameter
(parameter-shift) GPU gradient path, GPU-routed analytic post-processing,
gradients through the kernel-level drop-ins, and the backward of two states
that share one cached plan.  Each is checked against the complex128 oracle or
a closed form; tolerances: states 2e-6, energies 1e-6, gradients 1e-5."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import statevector as O
from tests.gpu_util import rand_state, rand_unitary

pytestmark = pytest.mark.gpu


# ---- driver payload of shots = 0 runs (driver.py:115-122) ----------------------

def test_driver_shots0_attaches_state_dispatch_for_cry(gpu):
    """state() skips cry (engine.py:914-1038) while run() applies it; the
    attached statevector must be state()'s, the expectations run()'s."""
    import tyxonq_amd as tq
    from tyxonq_amd import devices

    c = tq.Circuit(3).h(0).cry(0, 1, 1.1).rx(2, 0.4).measure_z(0).measure_z(1)
    (task,) = devices.run(provider="simulator", device="statevector", circuit=c, shots=0)
    r = task.get_result()["result_meta"]  # the driver's payload (driver.py:127-136)
    want_state, _ = O.simulate(3, c.ops, mode="state")
    np.testing.assert_allclose(r["statevector"], want_state, atol=2e-6)
    np.testing.assert_allclose(r["probabilities"], np.abs(want_state) ** 2, atol=2e-6)
    want_run, _ = O.simulate(3, c.ops, mode="run")
    assert r["expectations"]["Z1"] == pytest.approx(O.expect_z(want_run, 1, 3), abs=2e-6)
    assert abs(O.expect_z(want_run, 1, 3) - O.expect_z(want_state, 1, 3)) > 0.1


def test_driver_shots0_attaches_state_dispatch_for_inputs(gpu):
    """run() starts from |0..0> (engine.py:46); state() honours inputs (:906-912)."""
    import tyxonq_amd as tq
    from tyxonq_amd import devices

    rng = np.random.default_rng(3)
    init = rand_state(rng, 4)
    c = tq.Circuit(4, inputs=init).h(1).cx(1, 2).measure_z(2)
    (task,) = devices.run(provider="simulator", device="statevector", circuit=c, shots=0)
    r = task.get_result()["result_meta"]
    want_state, _ = O.simulate(4, c.ops, mode="state", init=init)
    np.testing.assert_allclose(r["statevector"], want_state, atol=2e-6)
    want_run, _ = O.simulate(4, c.ops, mode="run")
    assert r["expectations"]["Z2"] == pytest.approx(O.expect_z(want_run, 2, 4), abs=2e-6)


def test_driver_shots0_reuses_run_state_when_dispatches_agree(gpu):
    import tyxonq_amd as tq
    from tyxonq_amd import devices

    c = tq.Circuit(5).h(0).cx(0, 1).ry(3, 0.3).rzz(2, 4, 0.7).measure_z(1)
    (task,) = devices.run(provider="simulator", device="statevector", circuit=c, shots=0)
    r = task.get_result()["result_meta"]
    want, _ = O.simulate(5, c.ops, mode="state")
    np.testing.assert_allclose(r["statevector"], want, atol=2e-6)
    assert task.handle.device_state is not None and task.handle.device_state.is_cuda


# ---- depolarizing attenuation: only the op's wires (engine.py:56-141) ---------

def test_attenuation_wires_on_gpu_engine(gpu):
    from tyxonq_amd.circuit import Circuit
    from tyxonq_amd.engine import StatevectorEngine

    noise = {"type": "depolarizing", "p": 0.3}
    c = Circuit(3).rx(0, 1).rz(2, 5).cx(0, 1).measure_z(0).measure_z(1).measure_z(2)
    got = StatevectorEngine().run(c, shots=0, use_noise=True, noise=noise)["expectations"]
    want = O.run_expectations(3, c.ops, noise=noise)
    for k, v in want.items():
        assert got[k] == pytest.approx(v, abs=2e-6)
    assert got["Z0"] == pytest.approx(math.cos(1.0) * 0.36, abs=2e-6)
    assert got["Z2"] == pytest.approx(0.6, abs=2e-6)


# ---- perfect_sampling (engine.py:1061-1072) -----------------------------------

def test_perfect_sampling(gpu):
    import tyxonq_amd as tq
    from tyxonq_amd.engine import StatevectorEngine

    eng = StatevectorEngine()
    # a basis state: always that bitstring with probability 1
    bits, p = eng.perfect_sampling(tq.Circuit(4).x(1).x(3), rng=np.random.default_rng(0))
    assert bits == "0101" and p == pytest.approx(1.0, abs=1e-6)
    # GHZ: only 000 / 111, each with probability 1/2, both seen
    c = tq.Circuit(3).h(0).cx(0, 1).cx(1, 2)
    rng = np.random.default_rng(1)
    seen = set()
    for _ in range(40):
        b, pb = eng.perfect_sampling(c, rng=rng)
        assert b in ("000", "111") and pb == pytest.approx(0.5, abs=1e-6)
        seen.add(b)
    assert seen == {"000", "111"}
    # the returned probability is |amplitude|^2 of the returned bitstring
    c = tq.Circuit(4).ry(0, 0.7).ry(1, 1.9).cx(0, 2).rx(3, 2.2)
    want, _ = O.simulate(4, c.ops)
    for _ in range(10):
        b, pb = eng.perfect_sampling(c, rng=rng)
        assert pb == pytest.approx(abs(want[int(b, 2)]) ** 2, abs=2e-6)


# ---- the plain per-parameter GPU gradient path (common.py:11-23) --------------

def test_plain_parameter_shift_gpu_path(gpu):
    """kernels.parameter_shift_gradient over the GPU engine's energy (2P
    energies) equals the adjoint gradient and the oracle's parameter shift."""
    import torch

    import tyxonq_amd as tq
    from tyxonq_amd import bench_workloads as W
    from tyxonq_amd.engine import StatevectorEngine
    from tyxonq_amd.kernels import parameter_shift_gradient

    n, layers = 8, 2
    th0 = np.random.default_rng(5).normal(0, 0.5, size=n * layers)
    terms = W.tfim_terms(n, 0.8)
    eng = StatevectorEngine()

    def e_gpu(p):
        c, _ = W.hea_ry_cx(n, layers, params=list(p))
        return eng.expval(c, terms)

    g_plain = parameter_shift_gradient(e_gpu, th0)
    t = torch.tensor(th0, requires_grad=True)
    c, _ = W.hea_ry_cx(n, layers, params=[t[i] for i in range(t.numel())])
    _, g_adj = StatevectorEngine(backend_name="pytorch").energy_and_grad(c, terms)

    def e_ref(p):
        cc, _ = W.hea_ry_cx(n, layers, params=list(p))
        psi, _ = O.simulate(n, cc.ops)
        return O.pauli_sum_expectation(psi, terms, n)

    g_ref = O.parameter_shift_gradient(e_ref, th0)
    np.testing.assert_allclose(g_plain, g_ref, atol=1e-5)
    np.testing.assert_allclose(g_adj, g_ref, atol=1e-5)


# ---- shots = 0 Pauli sums reduced on the GPU (counts_expval.py:55-84) ---------

def test_shots0_pauli_sum_postprocessing_on_gpu(gpu):
    import tyxonq_amd as tq

    n = 6
    c = tq.Circuit(n).h(0).cx(0, 1).ry(2, 0.9).cx(2, 3).rx(4, 0.4).ry(5, 2.1)
    for q in range(n):
        c.measure_z(q)
    items = [(((0, "Z"), (1, "Z")), 0.7), (((2, "Z"),), -0.4), (((3, "Z"), (5, "Z")), 1.3),
             (((1, "Z"), (2, "Z"), (4, "Z")), 0.25), ((), 0.1)]
    c.postprocessing(method="expval_pauli_sum", items=items, identity_const=-0.5)
    res = c.run(shots=0)[0]
    post = res["postprocessing"]["result"]
    psi, _ = O.simulate(n, c.ops)
    p = np.abs(psi) ** 2
    want = []
    for term, _ in items:
        s = np.ones(1 << n)
        for q, _P in term:
            s *= 1 - 2 * ((np.arange(1 << n) >> (n - 1 - q)) & 1)
        want.append(float(np.dot(s, p)))
    np.testing.assert_allclose(post["expvals"], want, atol=2e-6)
    assert post["energy"] == pytest.approx(-0.5 + sum(cf * v for (_, cf), v in zip(items, want)),
                                           abs=5e-6)


# ---- gradients through the kernel-level drop-ins (statevector.py:28-68) -------

def _vqetfim_energy(K, G, param, n, nlayers, Jx, h):
    """The shape of examples/vqetfim_benchmark.py:70-103: RXX chain + RZ layers,
    h * sum <Z_i> + Jx * sum <X_i X_i+1> via H rotation and sign-weighted
    probabilities, through the kernel-level functions."""
    import torch

    psi = K.init_statevector(n)
    t = 0
    for _ in range(nlayers):
        for i in range(n - 1):
            psi = K.apply_2q_statevector(None, psi, G.gate_rxx(param[t, i]), i, i + 1, n)
        t += 1
        for i in range(n):
            psi = K.apply_1q_statevector(None, psi, G.gate_rz(param[t, i]), i, n)
        t += 1
    e = torch.zeros((), dtype=torch.float64)
    for i in range(n):
        e = e + h * K.expect_z_statevector(psi, i, n)
    psi_x = psi
    for q in range(n):
        psi_x = K.apply_1q_statevector(None, psi_x, G.gate_h(), q, n)
    probs = (psi_x.abs() ** 2).to(torch.float64)
    k = torch.arange(1 << n, device=probs.device)
    for i in range(n - 1):
        same = ((k >> (n - 1 - i)) & 1) == ((k >> (n - 2 - i)) & 1)
        e = e + Jx * torch.sum(torch.where(same, 1.0, -1.0).to(torch.float64) * probs)
    return e


def test_dropin_autograd_vqetfim_shape(gpu):
    import torch

    from tyxonq_amd import gates as G
    from tyxonq_amd import kernels as K

    n, nlayers, Jx, h = 6, 2, 1.0, -1.0
    p0 = np.random.default_rng(11).uniform(-1, 1, size=(2 * nlayers, n))
    p = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    e = _vqetfim_energy(K, G, p, n, nlayers, Jx, h)
    e.backward()
    terms = [(h, [("Z", i)]) for i in range(n)] + [(Jx, [("X", i), ("X", i + 1)])
                                                    for i in range(n - 1)]

    def ops_of(flat):
        pp = np.asarray(flat).reshape(2 * nlayers, n)
        ops = []
        for l in range(nlayers):
            ops += [("rxx", i, i + 1, pp[2 * l, i]) for i in range(n - 1)]
            ops += [("rz", i, pp[2 * l + 1, i]) for i in range(n)]
        return ops

    def e_ref(flat):
        psi, _ = O.simulate(n, ops_of(flat))
        return O.pauli_sum_expectation(psi, terms, n)

    assert float(e.detach()) == pytest.approx(e_ref(p0.reshape(-1)), abs=1e-6)
    g_ref = O.parameter_shift_gradient(e_ref, p0.reshape(-1)).reshape(p0.shape)
    # rxx rows of unused positions (index n-1) have zero gradient
    np.testing.assert_allclose(p.grad.numpy(), g_ref, atol=1e-5)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_dropin_autograd_state_and_matrix(gpu, k):
    """d/dpsi and d/dU of a real loss of U psi against torch-CPU complex128
    autograd of the oracle's contraction (statevector.py:71-129)."""
    import torch

    from tyxonq_amd import kernels as K

    n = 7
    rng = np.random.default_rng(20 + k)
    qubits = [5, 1, 3][:k]
    v0 = rand_state(rng, n)
    U0 = rand_unitary(rng, 1 << k)
    w = rng.normal(size=1 << n)
    a = rand_state(rng, n)

    def loss(out, wt, at):
        return torch.sum(wt * (out.abs() ** 2)) + torch.real(torch.sum(torch.conj(at) * out))

    # reference: torch CPU complex128 through the oracle's permutation contraction
    vs = torch.tensor(v0, requires_grad=True)
    Us = torch.tensor(U0, requires_grad=True)
    keep = [x for x in range(n) if x not in qubits]
    perm = keep + qubits
    inv = list(np.argsort(perm))
    t = vs.reshape((2,) * n).permute(perm).reshape(1 << (n - k), 1 << k)
    out_ref = (t @ Us.T).reshape((2,) * n).permute(inv).reshape(-1)
    loss(out_ref, torch.tensor(w), torch.tensor(a)).backward()

    vg = torch.tensor(v0, requires_grad=True)
    Ug = torch.tensor(U0, requires_grad=True)
    if k == 1:
        out = K.apply_1q_statevector(None, vg, Ug, qubits[0], n)
    elif k == 2:
        out = K.apply_2q_statevector(None, vg, Ug, qubits[0], qubits[1], n)
    else:
        out = K.apply_kqubit_unitary(vg, Ug, qubits, n)
    dev = out.device
    lg = loss(out.to(torch.complex128), torch.tensor(w, device=dev), torch.tensor(a, device=dev))
    lg.backward()
    assert vg.grad.dtype == torch.complex128 and vg.grad.device.type == "cpu"
    np.testing.assert_allclose(vg.grad.numpy(), vs.grad.numpy(), atol=2e-5)
    np.testing.assert_allclose(Ug.grad.numpy(), Us.grad.numpy(), atol=2e-5)


def test_dropin_expect_z_returns_tensor_and_grad(gpu):
    import torch

    from tyxonq_amd import kernels as K

    n = 5
    rng = np.random.default_rng(4)
    v0 = rand_state(rng, n)
    v = torch.tensor(v0, requires_grad=True)
    z = K.expect_z_statevector(v, 2, n)
    assert isinstance(z, torch.Tensor) and z.dim() == 0 and z.dtype == torch.float64
    assert float(z) == pytest.approx(O.expect_z(v0, 2, n), abs=1e-6)
    z.backward()
    vr = torch.tensor(v0, requires_grad=True)
    s = 1 - 2 * ((torch.arange(1 << n) >> (n - 1 - 2)) & 1)
    torch.sum(s * vr.abs() ** 2).backward()
    np.testing.assert_allclose(v.grad.numpy(), vr.grad.numpy(), atol=2e-6)
    with pytest.raises(ValueError):
        K.apply_1q_statevector_(v, np.eye(2), 0, n)


# ---- two states of one cached plan under one loss (engine.py _CircuitState) ---

def test_backward_of_two_states_sharing_a_plan(gpu):
    """psi_a = c(theta_a).state(), psi_b = c(theta_b).state(): the second
    lowering re-parameterises the cached plan; psi_a's backward must still use
    theta_a.  Loss = <Z0>_a + 2 <Z1>_b, against central differences."""
    import torch

    import tyxonq_amd as tq
    from tyxonq_amd import bench_workloads as W

    n, layers = 6, 2
    rng = np.random.default_rng(9)
    ta0 = rng.normal(0, 0.7, size=n * layers)
    tb0 = rng.normal(0, 0.7, size=n * layers)
    tq.set_backend("pytorch")
    try:
        ta = torch.tensor(ta0, requires_grad=True)
        tb = torch.tensor(tb0, requires_grad=True)
        ca, _ = W.hea_ry_cx(n, layers, params=[ta[i] for i in range(ta.numel())])
        cb, _ = W.hea_ry_cx(n, layers, params=[tb[i] for i in range(tb.numel())])
        psi_a = ca.state()
        psi_b = cb.state()
        za = 1 - 2 * ((torch.arange(1 << n, device=psi_a.device) >> (n - 1)) & 1)
        zb = 1 - 2 * ((torch.arange(1 << n, device=psi_a.device) >> (n - 2)) & 1)
        loss = torch.sum(za * psi_a.abs() ** 2) + 2 * torch.sum(zb * psi_b.abs() ** 2)
        loss.backward()
    finally:
        tq.set_backend("numpy")

    def z_of(th, q):
        c, _ = W.hea_ry_cx(n, layers, params=list(th))
        psi, _ = O.simulate(n, c.ops)
        return O.expect_z(psi, q, n)

    ga = O.central_fd_gradient(lambda t: z_of(t, 0), ta0, h=1e-4)
    gb = O.central_fd_gradient(lambda t: 2 * z_of(t, 1), tb0, h=1e-4)
    np.testing.assert_allclose(ta.grad.numpy(), ga, atol=1e-5)
    np.testing.assert_allclose(tb.grad.numpy(), gb, atol=1e-5)
