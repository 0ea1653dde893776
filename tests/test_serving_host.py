"""Serving request construction (CPU): the bit-level ops of a request equal
the oracle's simulation of the same RY/RZ/CX circuit (host emulator of the
pass kernel), and request generation is seeded."""
from __future__ import annotations

import numpy as np

from oracle import statevector as O
from tyxonq_amd import _native as nat
from tyxonq_amd import serving as S


def _oracle_ops(req):
    ops, k = [], 0
    for _ in range(req.depth):
        for q in range(req.n):
            ops.append(("ry", q, float(req.params[k])))
            k += 1
        for q in range(req.n):
            ops.append(("rz", q, float(req.params[k])))
            k += 1
        for q in range(req.n - 1):
            ops.append(("cx", q, q + 1))
    return ops


def test_request_ops_match_oracle():
    for req in S.make_requests(3, 100.0, seed=5, nmin=11, nmax=12, dmin=2, dmax=3):
        v = np.zeros(1 << req.n, dtype=np.complex64)
        v[0] = 1
        got = nat.emulate(req.n, S.request_ops(req), v)
        want, _ = O.simulate(req.n, _oracle_ops(req))
        np.testing.assert_allclose(got, want, atol=2e-5)


def test_requests_are_seeded_and_in_range():
    a = S.make_requests(50, 200.0, seed=9)
    b = S.make_requests(50, 200.0, seed=9)
    assert [r.n for r in a] == [r.n for r in b]
    assert all(10 <= r.n <= 26 and 2 <= r.depth <= 16 for r in a)
    assert all(a[i].arrival < a[i + 1].arrival for i in range(49))
    assert all(len(r.params) == 2 * r.n * r.depth for r in a)
    assert S.Server.footprint(a[0]) == 24.0 * (1 << a[0].n)


def test_request_coeffs_follow_terms():
    for req in S.make_requests(6, 100.0, seed=2, nmin=5, nmax=8):
        np.testing.assert_array_equal(S.request_coeffs(req), [c for c, _ in S.request_terms(req)])
