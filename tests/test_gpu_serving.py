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
