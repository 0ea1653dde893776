"""Pins the CPU oracle (CPU only).

The oracle is checked against (a) the reference's own known answers
(tests/golden/reference_known_answers.json), (b) an independent dense
Kronecker brute force (oracle/dense.py) and (c) closed forms, so that a
dropped term, a wrong sign, a swapped bit order or a transposed two-qubit
matrix fails at least one of them.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import dense
from oracle import statevector as O
from tests.conftest import cplx, load_known_answers

KA = load_known_answers()


def ka_ops(entry, params=None):
    ops = []
    for op in entry.get("ops", []):
        op = list(op)
        if params is not None:
            op = [params[int(a[1:])] if isinstance(a, str) and a.startswith("p") and a[1:].isdigit()
                  else a for a in op]
        ops.append(tuple(op))
    ucache = {k: cplx(v) for k, v in entry.get("unitaries", {}).items()}
    return ops, ucache


@pytest.mark.parametrize("entry", KA, ids=[e["name"] for e in KA])
def test_oracle_known_answers(entry):
    n = entry.get("n")
    atol = entry.get("atol", 1e-10)
    if "matrix" in entry:
        want = cplx(entry["matrix"])
        if "codes" in entry:
            got = dense.pauli_string([("IXYZ"[c], q) for q, c in enumerate(entry["codes"][0])
                                      if c], len(entry["codes"][0]))
        else:
            got = O.build_controlled_unitary(cplx(entry["controlled"]["U"]),
                                             entry["controlled"]["num_controls"])
        np.testing.assert_allclose(got, want, atol=1e-12)
        return
    if "run_expectations" in entry:
        ops, uc = ka_ops(entry)
        got = O.run_expectations(n, ops, noise=entry.get("noise"), unitary_cache=uc)
        for k, v in entry["run_expectations"].items():
            assert abs(got[k] - v) < atol
        return
    if "grad" in entry:
        terms = [(c, [tuple(f) for f in fac]) for c, fac in entry["pauli_terms"]]

        def energy(p):
            ops, _ = ka_ops(entry, p)
            psi, _ = O.simulate(n, ops)
            return O.pauli_sum_expectation(psi, terms, n)

        g = O.parameter_shift_gradient(energy, entry["params"])
        np.testing.assert_allclose(g, entry["grad"], atol=atol)
        return
    ops, uc = ka_ops(entry)
    psi, _ = O.simulate(n, ops, unitary_cache=uc)
    if "state" in entry:
        np.testing.assert_allclose(psi, cplx(entry["state"]), atol=atol)
    if "z" in entry:
        for q, v in entry["z"].items():
            assert abs(O.expect_z(psi, int(q), n) - v) < atol
    if "norm" in entry:
        assert abs(np.vdot(psi, psi).real - entry["norm"]) < atol
    if "prob_sum_indices" in entry:
        idx, val = entry["prob_sum_indices"]
        assert abs(np.sum(np.abs(psi[idx]) ** 2) - val) < atol
    if "amplitudes" in entry:
        for bits, v in entry["amplitudes"].items():
            assert abs(abs(psi[int(bits, 2)]) - v) < atol
    if "pauli_terms" in entry:
        terms = [(c, [tuple(f) for f in fac]) for c, fac in entry["pauli_terms"]]
        assert abs(O.pauli_sum_expectation(psi, terms, n) - entry["energy"]) < atol


# ---- random circuits vs. the dense brute force ------------------------------

_ONE = ["h", "x", "s", "sdg", "rx", "ry", "rz"]
_TWO = ["cx", "cz", "iswap", "swap", "rxx", "ryy", "rzz", "cry"]


def random_ops(rng, n, depth, two=True):
    ops = []
    for _ in range(depth):
        if n >= 2 and two and rng.random() < 0.4:
            a, b = rng.choice(n, 2, replace=False)
            name = _TWO[rng.integers(len(_TWO))]
            if name in ("rxx", "ryy", "rzz", "cry"):
                ops.append((name, int(a), int(b), float(rng.uniform(0, 2 * np.pi))))
            else:
                ops.append((name, int(a), int(b)))
        else:
            name = _ONE[rng.integers(len(_ONE))]
            q = int(rng.integers(n))
            if name in ("rx", "ry", "rz"):
                ops.append((name, q, float(rng.uniform(0, 2 * np.pi))))
            else:
                ops.append((name, q))
    return ops


def dense_simulate(n, ops):
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1
    one = {"h": O.gate_h(), "x": O.X, "s": O.gate_s(), "sdg": O.gate_sd()}
    for op in ops:
        name = op[0]
        if name in one:
            psi = dense.op_1q(one[name], op[1], n) @ psi
        elif name in ("rx", "ry", "rz"):
            P = {"rx": O.X, "ry": O.Y, "rz": O.Z}[name]
            t = op[2]
            U = math.cos(t / 2) * np.eye(2) - 1j * math.sin(t / 2) * P
            psi = dense.op_1q(U, op[1], n) @ psi
        else:
            U = {"cx": O.gate_cx, "cz": O.gate_cz, "iswap": O.gate_iswap, "swap": O.gate_swap}.get(name)
            if U is not None:
                U = U()
            elif name == "cry":
                U = np.eye(4, dtype=complex)
                t = op[3]
                U[2:, 2:] = [[math.cos(t / 2), -math.sin(t / 2)], [math.sin(t / 2), math.cos(t / 2)]]
            else:
                P = {"rxx": O.X, "ryy": O.Y, "rzz": O.Z}[name]
                t = op[3]
                U = math.cos(t / 2) * np.eye(4) - 1j * math.sin(t / 2) * np.kron(P, P)
            psi = dense.op_2q(U, op[1], op[2], n) @ psi
    return psi


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_oracle_matches_dense_bruteforce(n):
    rng = np.random.default_rng(100 + n)
    ops = random_ops(rng, n, 40)
    # run() dispatches cry, state() does not: compare the run-mode state
    psi, _ = O.simulate(n, ops, mode="run")
    np.testing.assert_allclose(psi, dense_simulate(n, ops), atol=1e-12)


def test_state_mode_skips_reference_ignored_ops():
    n = 2
    ops = [("h", 0), ("cry", 0, 1, 0.7), ("z", 0), ("y", 1), ("t", 0)]
    psi, _ = O.simulate(n, ops, mode="state")
    np.testing.assert_allclose(psi, dense_simulate(n, [("h", 0)]), atol=1e-14)
    psi_all, _ = O.simulate(n, ops, mode="state", reference_dispatch=False)
    assert not np.allclose(psi_all, psi)


def test_apply_kqubit_matches_apply_2q():
    rng = np.random.default_rng(7)
    n = 5
    psi = rng.normal(size=32) + 1j * rng.normal(size=32)
    U = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
    for q0, q1 in [(0, 1), (3, 1), (4, 0), (2, 3)]:
        np.testing.assert_allclose(O.apply_kqubit(psi, U, [q0, q1], n), O.apply_2q(psi, U, q0, q1, n),
                                   atol=1e-12)
        np.testing.assert_allclose(O.apply_2q(psi, U, q0, q1, n), dense.op_2q(U, q0, q1, n) @ psi,
                                   atol=1e-12)


def random_terms(rng, n, k):
    terms = []
    for _ in range(k):
        w = int(rng.integers(1, min(n, 4) + 1))
        qs = rng.choice(n, w, replace=False)
        terms.append((float(rng.normal()), [("XYZ"[int(rng.integers(3))], int(q)) for q in qs]))
    return terms


@pytest.mark.parametrize("n", [1, 3, 5, 7])
def test_pauli_sum_oracle_vs_dense(n):
    rng = np.random.default_rng(n)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    terms = random_terms(rng, n, 25)
    e = O.pauli_sum_expectation(psi, terms, n)
    H = dense.pauli_sum_dense(terms, n)
    assert abs(e - np.real(np.conj(psi) @ H @ psi)) < 1e-12
    for q in range(n):
        assert abs(O.expect_z(psi, q, n) - dense.z_expect(psi, q, n)) < 1e-12


def test_expectation_product_matches_dense():
    rng = np.random.default_rng(3)
    n = 4
    psi = rng.normal(size=16) + 1j * rng.normal(size=16)
    psi /= np.linalg.norm(psi)
    got = O.expectation_product(psi, [(O.X, [1]), (O.Z, [3])], n)
    want = np.real(np.conj(psi) @ dense.pauli_string([("X", 1), ("Z", 3)], n) @ psi)
    assert abs(got - want) < 1e-12


# ---- closed forms -------------------------------------------------------------

@pytest.mark.parametrize("t", [0.0, 0.3, 1.7, math.pi, 5.1])
def test_rx_z_expectation_closed_form(t):
    psi, _ = O.simulate(1, [("rx", 0, t)])
    assert abs(O.expect_z(psi, 0, 1) - math.cos(t)) < 1e-13


def test_rotation_identities():
    a, b = 0.37, 1.21
    np.testing.assert_allclose(O.gate_rz(a) @ O.gate_rz(b), O.gate_rz(a + b), atol=1e-14)
    np.testing.assert_allclose(O.gate_rx(2 * math.pi), -np.eye(2), atol=1e-14)
    np.testing.assert_allclose(O.gate_ry(a), math.cos(a / 2) * np.eye(2) - 1j * math.sin(a / 2) * O.Y,
                               atol=1e-14)


def _evolve_pauli_gates(paulis, theta):
    """The gate sequence of evolve_pauli_ops (libs/circuits_library/utils.py:29-59)."""
    ops = []
    for q, P in paulis:
        if P == "X":
            ops.append(("h", q))
        elif P == "Y":
            ops += [("sdg", q), ("h", q)]
    for i in range(len(paulis) - 1):
        ops.append(("cx", paulis[i][0], paulis[i + 1][0]))
    ops.append(("rz", paulis[-1][0], theta))
    for i in reversed(range(len(paulis) - 1)):
        ops.append(("cx", paulis[i][0], paulis[i + 1][0]))
    for q, P in paulis:
        if P == "X":
            ops.append(("h", q))
        elif P == "Y":
            ops += [("h", q), ("s", q)]
    return ops


@pytest.mark.parametrize("paulis", [[(0, "Z")], [(1, "X"), (3, "Y")], [(0, "Y"), (1, "Z"), (2, "X")],
                                    [(0, "X"), (1, "Y"), (2, "Y"), (3, "X")]])
def test_pauli_rotation_matches_reference_gate_sequence(paulis):
    n = 4
    rng = np.random.default_rng(11)
    psi = rng.normal(size=16) + 1j * rng.normal(size=16)
    psi /= np.linalg.norm(psi)
    th = 0.813
    got = O.pauli_rotation(psi, [(P, q) for q, P in paulis], th, n)
    want, _ = O.simulate(n, _evolve_pauli_gates(paulis, th), init=psi)
    np.testing.assert_allclose(got, want, atol=1e-12)


def _hea(th, n, layers):
    ops = []
    k = 0
    for _ in range(layers):
        for q in range(n):
            ops.append(("ry", q, th[k]))
            k += 1
        for q in range(n):
            ops.append(("rz", q, th[k]))
            k += 1
        for q in range(n - 1):
            ops.append(("cx", q, q + 1))
    return ops


def test_gradient_rules_agree():
    n, layers = 4, 2
    rng = np.random.default_rng(5)
    th = rng.normal(0, 0.7, size=2 * n * layers)
    terms = random_terms(rng, n, 12)

    def energy(p):
        psi, _ = O.simulate(n, _hea(p, n, layers))
        return O.pauli_sum_expectation(psi, terms, n)

    ps = O.parameter_shift_gradient(energy, th)
    fd = O.central_fd_gradient(energy, th, h=1e-4)
    e_ag, ag = O.torch_autograd_energy(n, lambda t: _hea(t, n, layers), terms, th)
    assert abs(e_ag - energy(th)) < 1e-12
    np.testing.assert_allclose(ps, ag, atol=1e-10)
    np.testing.assert_allclose(ps, fd, atol=1e-7)


# ---- Kraus trajectory (statevector.py:132-218) pinned by closed forms ---------

def _amp_damp(g):
    return [np.array([[1, 0], [0, math.sqrt(1 - g)]], dtype=complex),
            np.array([[0, math.sqrt(g)], [0, 0]], dtype=complex)]


@pytest.mark.parametrize("g", [0.1, 0.35, 0.8])
def test_kraus_trajectory_amplitude_damping_on_one(g):
    """|1>: Born weights (1-g, g); status <= 1-g keeps |1>, above it decays to |0>."""
    one = np.array([0, 1], dtype=complex)
    below = O.apply_kraus(one, _amp_damp(g), 0, 1, status=(1 - g) * 0.999)
    np.testing.assert_allclose(below, [0, 1], atol=1e-14)
    above = O.apply_kraus(one, _amp_damp(g), 0, 1, status=(1 - g) + 0.5 * g)
    np.testing.assert_allclose(above, [1, 0], atol=1e-14)


@pytest.mark.parametrize("g", [0.2, 0.6])
def test_kraus_trajectory_bell_state_closed_form(g):
    """(|00>+|11>)/sqrt2 with damping on qubit 1 (index bit 0): K0 is chosen with
    probability 1-g/2 and gives (|00> + sqrt(1-g)|11>)/sqrt(2-g); K1 gives |10>
    (index 2).  A wrong qubit axis, a dropped renormalisation or a selection
    against the unnormalised weights fails one of the four checks."""
    bell = np.array([1, 0, 0, 1], dtype=complex) / math.sqrt(2)
    p0 = 1 - g / 2
    k0 = O.apply_kraus(bell, _amp_damp(g), 1, 2, status=p0 - 1e-9)
    want0 = np.array([1, 0, 0, math.sqrt(1 - g)], dtype=complex) / math.sqrt(2 - g)
    np.testing.assert_allclose(k0, want0, atol=1e-14)
    k1 = O.apply_kraus(bell, _amp_damp(g), 1, 2, status=p0 + 1e-9)
    np.testing.assert_allclose(k1, [0, 0, 1, 0], atol=1e-14)
    # damping on qubit 0 (index bit 1) instead: K1 gives |01> (index 1)
    k1q0 = O.apply_kraus(bell, _amp_damp(g), 0, 2, status=p0 + 1e-9)
    np.testing.assert_allclose(k1q0, [0, 1, 0, 0], atol=1e-14)
    # status beyond every cumulative weight falls back to K0 (selected index 0)
    np.testing.assert_allclose(O.apply_kraus(bell, _amp_damp(g), 1, 2, status=1.5), want0,
                               atol=1e-14)


def test_kraus_trajectory_unnormalised_kraus_set():
    """Weights are normalised before selection (statevector.py:189-192): K0 = I/2,
    K1 = X/2 on |0> have weights 1/4 and 1/4, so status 0.49 picks K0 and 0.51 K1;
    the result is renormalised either way."""
    ks = [0.5 * np.eye(2, dtype=complex), 0.5 * O.X]
    zero = np.array([1, 0], dtype=complex)
    np.testing.assert_allclose(O.apply_kraus(zero, ks, 0, 1, status=0.49), [1, 0], atol=1e-14)
    np.testing.assert_allclose(O.apply_kraus(zero, ks, 0, 1, status=0.51), [0, 1], atol=1e-14)


def test_kraus_docstring_example():
    """The reference's own example (statevector.py:168-171): amplitude damping
    gamma=0.1 on qubit 0 of |00>, status 0.5 -> K0 (weight 1) -> |00>."""
    zero = np.zeros(4, dtype=complex)
    zero[0] = 1
    np.testing.assert_allclose(O.apply_kraus(zero, _amp_damp(0.1), 0, 2, status=0.5), zero,
                               atol=1e-14)


# ---- depolarizing attenuation touches only an op's wires (engine.py:56-141) ----

def test_attenuation_ignores_integer_angles():
    """rx(0, 1): the angle 1 is not a wire; <Z1> is not attenuated."""
    noise = {"type": "depolarizing", "p": 0.3}
    got = O.run_expectations(2, [("rx", 0, 1), ("measure_z", 0), ("measure_z", 1)], noise=noise)
    assert abs(got["Z0"] - math.cos(1.0) * 0.6) < 1e-12
    assert abs(got["Z1"] - 1.0) < 1e-12
    # an integer angle beyond the register must not index a qubit
    got = O.run_expectations(3, [("rz", 2, 5), ("cx", 0, 1), ("measure_z", 0), ("measure_z", 2)],
                             noise=noise)
    assert abs(got["Z0"] - 0.6) < 1e-12 and abs(got["Z2"] - 0.6) < 1e-12
    # a unitary without a cached matrix is skipped and attenuates nothing
    got = O.run_expectations(1, [("unitary", 0, "missing"), ("measure_z", 0)], noise=noise)
    assert abs(got["Z0"] - 1.0) < 1e-12
