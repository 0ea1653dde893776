"""Serving request construction (CPU): the bit-level ops of a request equal
the oracle's simulation of the same RY/RZ/CX circuit (host emulator of the
pass kernel), and request generation is seeded."""
from __future__ import annotations

import numpy as np
import pytest

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


# ---- server logic with stand-in executors (CPU) --------------------------------

def _fake_exec(fail_ids=(), delay=0.0):
    import time

    def run(req, cache):
        if delay:
            time.sleep(delay)
        if req.rid in fail_ids:
            raise RuntimeError(f"forced failure of request {req.rid}")
        return float(req.n), np.full(3, float(req.rid))

    def run_batch(reqs, cache):
        return [run(r, cache) for r in reqs]

    return run, run_batch


def test_server_isolates_failures_and_rejects_never_fitting():
    reqs = S.make_requests(30, 1e6, seed=11, nmin=10, nmax=22, dmin=2, dmax=3)
    small = [r.rid for r in reqs if S.is_small(r)]
    large = [r.rid for r in reqs if not S.is_small(r)]
    fail_ids = {small[0], large[0]}
    run, run_batch = _fake_exec(fail_ids)
    budget = 24.0 * (1 << 20) + 1  # 21+ qubit requests never fit
    srv = S.Server(slots=4, mem_budget=budget, execute_fn=run, execute_batch_fn=run_batch,
                   streams=False)
    res = srv.run(reqs, realtime=False)
    assert [r.rid for r in res] == [r.rid for r in reqs]
    for req, r in zip(reqs, res):
        if req.rid in fail_ids:
            assert not r.ok and "forced failure" in r.error
        elif S.Server.footprint(req) > budget:
            assert not r.ok and r.error.startswith("rejected")
        else:
            assert r.ok and r.energy == float(req.n) and r.grad[0] == req.rid
    assert srv.last_stats["rejected"] == sum(S.Server.footprint(r) > budget for r in reqs)
    assert srv.last_stats["failed"] == 2
    summ = S.summarize(res)
    assert summ["failed"] == 2 + srv.last_stats["rejected"] and summ["p99_ms"] >= 0


def test_server_share_admission_runs_chip_filling_requests_alone():
    import threading
    import time

    active = {"now": 0.0, "max": 0.0}
    lock = threading.Lock()

    def run(req, cache):
        with lock:
            active["now"] += S.gpu_share(req.n)
            active["max"] = max(active["max"], active["now"])
        time.sleep(0.01)
        with lock:
            active["now"] -= S.gpu_share(req.n)
        return 0.0, np.zeros(1)

    reqs = S.make_requests(24, 1e6, seed=12, nmin=14, nmax=26, dmin=2, dmax=3)
    S.Server(slots=8, execute_fn=run, execute_batch_fn=lambda rs, c: [run(r, c) for r in rs],
             streams=False).run(reqs, realtime=False)
    assert active["max"] <= 2.0 + 1e-9
    assert S.gpu_share(22) == 1.0 and S.gpu_share(13) == pytest.approx(1 / 512)


def test_validate_rejects_bad_requests():
    req = S.make_requests(1, 10.0, seed=1, nmin=10, nmax=10, dmin=2, dmax=2)[0]
    bad = S.Request(req.rid + 100, req.n, req.depth, req.ham, req.g, req.params.copy(), req.arrival)
    bad.params[3] = np.nan
    with pytest.raises(ValueError, match="non-finite"):
        S.request_ops(bad)
    short = S.Request(req.rid, req.n, req.depth, req.ham, req.g, req.params[:-1], req.arrival)
    with pytest.raises(ValueError, match="angles"):
        S.request_ops(short)
    res = S.plain_loop([bad, req], realtime=False, execute_fn=lambda r, c: (S.validate(r), (1.0, np.zeros(1)))[1],
                       sync=False)
    assert not res[0].ok and res[1].ok
