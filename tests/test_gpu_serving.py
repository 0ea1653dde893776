"""Continuous-batching server (GPU): every request computes the same energy
and gradient alone and in the mix (bit-identical: fixed-order reductions,
one stream per request), and a request's energy matches the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import statevector as O
from tyxonq_amd import serving as S

pytestmark = pytest.mark.gpu


def test_results_alone_and_in_mix(gpu):
    reqs = S.make_requests(24, 1e6, seed=3, nmin=10, nmax=15, dmin=2, dmax=6)
    mixed = S.Server(slots=6, mem_budget=1e9).run(reqs, realtime=False)
    alone = S.plain_loop(reqs, realtime=False)
    for a, m in zip(alone, mixed):
        assert a.rid == m.rid
        assert a.energy == m.energy
        np.testing.assert_array_equal(a.grad, m.grad)


def test_server_admission_respects_budget(gpu):
    # a budget for about two 15-qubit footprints: still completes every request
    reqs = S.make_requests(8, 1e6, seed=4, nmin=14, nmax=15, dmin=2, dmax=3)
    res = S.Server(slots=8, mem_budget=2.2 * 24 * (1 << 15)).run(reqs, realtime=False)
    assert [r.rid for r in res] == [r.rid for r in reqs]
    assert all(np.isfinite(r.energy) for r in res)


def test_request_energy_matches_oracle(gpu):
    req = S.make_requests(1, 10.0, seed=6, nmin=10, nmax=10, dmin=3, dmax=3)[0]
    (res,) = S.plain_loop([req], realtime=False)
    ops = []
    k = 0
    for _ in range(req.depth):
        for q in range(req.n):
            ops.append(("ry", q, float(req.params[k])))
            k += 1
        for q in range(req.n):
            ops.append(("rz", q, float(req.params[k])))
            k += 1
        for q in range(req.n - 1):
            ops.append(("cx", q, q + 1))
    psi, _ = O.simulate(req.n, ops)
    e = O.pauli_sum_expectation(psi, S.request_terms(req), req.n)
    assert abs(res.energy - e) < 1e-6

    def f(p):
        o, j = [], 0
        for op in ops:
            if op[0] in ("ry", "rz"):
                o.append((op[0], op[1], float(p[j])))
                j += 1
            else:
                o.append(op)
        s, _ = O.simulate(req.n, o)
        return O.pauli_sum_expectation(s, S.request_terms(req), req.n)

    idx = [0, 7, 13, len(req.params) - 1]
    for j in idx:
        d = np.zeros(len(req.params))
        d[j] = np.pi / 2
        ps = 0.5 * (f(req.params + d) - f(req.params - d))
        assert abs(res.grad[j] - ps) < 1e-5


def _oracle_energy_grad(req, idx):
    ops = []
    k = 0
    for _ in range(req.depth):
        for q in range(req.n):
            ops.append(("ry", q, float(req.params[k])))
            k += 1
        for q in range(req.n):
            ops.append(("rz", q, float(req.params[k])))
            k += 1
        for q in range(req.n - 1):
            ops.append(("cx", q, q + 1))

    def f(p):
        o, j = [], 0
        for op in ops:
            if op[0] in ("ry", "rz"):
                o.append((op[0], op[1], float(p[j])))
                j += 1
            else:
                o.append(op)
        s, _ = O.simulate(req.n, o)
        return O.pauli_sum_expectation(s, S.request_terms(req), req.n)

    e = f(req.params)
    g = {}
    for j in idx:
        d = np.zeros(len(req.params))
        d[j] = np.pi / 2
        g[j] = 0.5 * (f(req.params + d) - f(req.params - d))
    return e, g


def _batch_inputs(reqs):
    from tyxonq_amd import _native as nat
    from tyxonq_amd.pauli import PauliSum

    plans, hams, psis = [], [], []
    for r in reqs:
        plans.append(nat.Plan(r.n, S.request_ops(r), min_qubits=13, jit=-1))
        x, z, c = PauliSum(r.n, S.request_terms(r)).to_masks()
        hams.append(nat.Ham(r.n, x, z, c))
        psis.append(nat.new_state(r.n, 0))
    return plans, hams, psis


def test_batched_launch_matches_oracle_and_ignores_the_mix(gpu):
    # requests of 2..13 qubits and depths 1..5 in one launch: each energy and
    # gradient matches the complex128 oracle, and a request's result is the
    # same bit for bit in a batch of any composition
    from tyxonq_amd import _native as nat

    reqs = S.make_requests(12, 1e6, seed=9, nmin=2, nmax=13, dmin=1, dmax=5)
    plans, hams, psis = _batch_inputs(reqs)
    batch = nat.Batch()
    full = batch.adjoint(plans, hams, psis, [len(r.params) for r in reqs])
    for r, (e, g) in zip(reqs, full):
        idx = sorted({0, len(r.params) // 2, len(r.params) - 1})
        we, wg = _oracle_energy_grad(r, idx)
        assert abs(e - we) < 1e-6, (r.n, r.depth)
        for j in idx:
            assert abs(g[j] - wg[j]) < 1e-5, (r.n, r.depth, j)
    sub = [3, 7, 0]
    part = batch.adjoint([plans[i] for i in sub], [hams[i] for i in sub], [psis[i] for i in sub],
                         [len(reqs[i].params) for i in sub])
    for k, i in enumerate(sub):
        assert part[k][0] == full[i][0]
        np.testing.assert_array_equal(part[k][1], full[i][1])
    # the single-request path (separate launches) agrees to rounding
    for r, p, h, (e, g) in zip(reqs, plans, hams, full):
        psi = nat.new_state(r.n, 0)
        lam = nat.new_state(r.n, 0)
        e1, g1 = nat.Plan(r.n, S.request_ops(r), jit=-1).adjoint(h, psi, lam, len(r.params))
        assert abs(e1 - e) < 1e-6
        np.testing.assert_allclose(g1, g, atol=1e-5)


def test_batch_rejects_plans_not_built_for_it(gpu):
    from tyxonq_amd import _native as nat

    reqs = S.make_requests(2, 1e6, seed=2, nmin=5, nmax=6, dmin=1, dmax=2)
    plans, hams, psis = _batch_inputs(reqs)
    plans[1] = nat.Plan(reqs[1].n, S.request_ops(reqs[1]))  # built for 11 qubits
    with pytest.raises(ValueError):
        nat.Batch().adjoint(plans, hams, psis, [len(r.params) for r in reqs])


def test_failed_request_is_isolated_and_others_unchanged(gpu):
    """A request with a non-finite angle fails alone (host-side validation before
    any launch); every other request equals its result alone.  A request whose
    footprint can never fit the budget is rejected at arrival."""
    reqs = S.make_requests(12, 1e6, seed=21, nmin=10, nmax=16, dmin=2, dmax=4)
    bad_small = next(r for r in reqs if S.is_small(r))
    bad_large = next(r for r in reqs if not S.is_small(r))
    bad_small.params[0] = np.nan
    bad_large.params[-1] = np.inf
    huge = S.Request(999, 18, 2, "tfim", 1.0, np.zeros(2 * 18 * 2), 0.0)
    budget = 24.0 * (1 << 15) * 4
    mixed = S.Server(slots=4, mem_budget=budget).run(reqs + [huge], realtime=False)
    good = [r for r in reqs if r.rid not in (bad_small.rid, bad_large.rid)]
    alone = {r.rid: r for r in S.plain_loop(good, realtime=False)}
    for req, res in zip(reqs + [huge], mixed):
        if req.rid in (bad_small.rid, bad_large.rid):
            assert not res.ok and "non-finite" in res.error
        elif req.rid == 999 or S.Server.footprint(req) > budget:
            assert not res.ok and res.error.startswith("rejected")
        else:
            assert res.ok and res.energy == alone[req.rid].energy
            np.testing.assert_array_equal(res.grad, alone[req.rid].grad)
