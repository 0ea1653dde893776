"""Energy+gradient serving with continuous batching on one GPU.

Requests (qubit count, hardware-efficient ansatz depth and angles, TFIM or
Heisenberg Hamiltonian) arrive over time.  Requests of up to 13 qubits are
served in batched launches: a batch lane takes every small request that has
arrived and computes all their energies and gradients in one kernel launch,
one work-group per request with its state and lambda in LDS
(tqsv_batch_adjoint).  Larger requests are admitted in arrival order as soon
as they have arrived, a slot is free, their adjoint footprint (2^n x 24 B: the
state, lambda and headroom) fits the memory budget and their share of the
GPU fits beside the requests already running.  The share of a request is the
fraction of the chip its tile passes fill: 2^(n-13) work-groups of 2^13
amplitudes against 512 resident work-groups (2 per CU), capped at 1; the
shares in flight may sum to `share_cap` (default 2): two chip-filling
requests overlap each other's host work and synchronisations, a third would
only time-slice the GPU and lengthen all of them (measured on the serving
workload at 50, 200 and 1000 req/s offered: cap 2 gave the best p50, p99 and
throughput of caps 1, 2 and unlimited).  Every admitted request runs on a worker with its own
HIP stream and resident device states.  When one finishes, its slot, memory
and share are released and the next request is admitted.  A request whose
footprint can never fit the budget is rejected at arrival, and a request that
fails (invalid parameters, an error from the library) gets its error in its
result while the others proceed.  `plain_loop` runs the same requests one at
a time in arrival order on one stream, for comparison.

Each request's result depends only on the request: the reductions have a fixed
order, so a request computes the same energy and gradient alone and in any mix.
"""
from __future__ import annotations

import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from functools import lru_cache
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _native as nat
from .pauli import PauliSum

Terms = List[Tuple[float, List[Tuple[str, int]]]]


@dataclass
class Request:
    rid: int
    n: int
    depth: int
    ham: str            # "tfim" or "heisenberg"
    g: float
    params: np.ndarray  # 2 * n * depth angles: RY then RZ per layer
    arrival: float = 0.0


@dataclass
class Result:
    rid: int
    energy: float
    grad: np.ndarray
    arrival: float
    start: float
    done: float
    error: Optional[str] = None  # set when the request failed or was rejected

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    @property
    def ok(self) -> bool:
        return self.error is None


def make_requests(count: int, rate: float, seed: int = 1234, nmin: int = 10, nmax: int = 26,
                  dmin: int = 2, dmax: int = 16) -> List[Request]:
    """Poisson arrivals at `rate` per second; per request n ~ U{nmin..nmax},
    depth ~ U{dmin..dmax}, H in {TFIM (open chain), Heisenberg (periodic)}
    with g ~ U[0.5, 1.5], angles ~ U[0, 2pi).  Draw order per request: gap,
    n, depth, Hamiltonian kind, g, angles."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(rng.integers(nmin, nmax + 1))
        d = int(rng.integers(dmin, dmax + 1))
        kind = "tfim" if rng.random() < 0.5 else "heisenberg"
        g = float(rng.uniform(0.5, 1.5))
        th = rng.uniform(0.0, 2.0 * math.pi, size=2 * n * d)
        out.append(Request(k, n, d, kind, g, th, t))
    return out


def request_terms(req: Request) -> Terms:
    n, g = req.n, req.g
    if req.ham == "tfim":
        t: Terms = [(-1.0, [("Z", i), ("Z", i + 1)]) for i in range(n - 1)]
        return t + [(-g, [("X", i)]) for i in range(n)]
    t = []
    for i in range(n):
        j = (i + 1) % n
        for P in "XYZ":
            t.append((1.0, [(P, i), (P, j)]))
    return t + [(g, [("Z", i)]) for i in range(n)]


@lru_cache(maxsize=None)
def _ops_template(n: int, depth: int):
    """Op rows of the ansatz with zero angles, and the row indices of the RY
    and RZ rotations (parameter k is row ry[k] or rz[k])."""
    rows, ry, rz = [], [], []
    X = [0.0] * 32
    X[2] = X[4] = 1.0
    k = 0
    for _ in range(depth):
        for q in range(n):
            b = 1 << (n - 1 - q)
            ry.append(len(rows))
            rows.append((nat.OP_MAT1, (n - 1 - q, -1), k, 0, 0, 0, 0, 0.0, b, b, 1.0, [0.0] * 32))
            k += 1
        for q in range(n):
            b = 1 << (n - 1 - q)
            rz.append(len(rows))
            rows.append((nat.OP_MAT1, (n - 1 - q, -1), k, 0, 0, 0, 0, 0.0, 0, b, 1.0, [0.0] * 32))
            k += 1
        for q in range(n - 1):
            cb = 1 << (n - 1 - q)
            rows.append((nat.OP_MAT1, (n - 2 - q, -1), -1, cb, cb, 0, 0, 0.0, 0, 0, 0.0, X))
    return np.array(rows, dtype=nat.OP_DTYPE), np.array(ry), np.array(rz)


def validate(req: Request) -> None:
    """Host-side checks before anything is launched for a request."""
    if not (1 <= int(req.n) <= 33):
        raise ValueError(f"request {req.rid}: {req.n} qubits is outside [1, 33]")
    if int(req.depth) < 1:
        raise ValueError(f"request {req.rid}: depth must be positive")
    th = np.asarray(req.params, dtype=np.float64)
    if th.size != 2 * req.n * req.depth:
        raise ValueError(f"request {req.rid}: expected {2 * req.n * req.depth} angles, "
                         f"got {th.size}")
    if not np.all(np.isfinite(th)) or not math.isfinite(req.g):
        raise ValueError(f"request {req.rid}: non-finite angle or field")


def request_ops(req: Request) -> np.ndarray:
    """depth x [RY(theta) on every qubit, RZ(theta) on every qubit, CX ladder
    q -> q+1], as bit-level ops with one parameter per rotation."""
    validate(req)
    tmpl, ry, rz = _ops_template(req.n, req.depth)
    ops = tmpl.copy()
    th = np.asarray(req.params, dtype=np.float64).reshape(req.depth, 2, req.n)
    c, s = np.cos(th / 2), np.sin(th / 2)
    m = ops["m"]
    # RY = [[c, -s], [s, c]];  RZ = diag(c - i s, c + i s)
    m[ry, 0] = c[:, 0].ravel()
    m[ry, 2] = -s[:, 0].ravel()
    m[ry, 4] = s[:, 0].ravel()
    m[ry, 6] = c[:, 0].ravel()
    m[rz, 0] = c[:, 1].ravel()
    m[rz, 1] = -s[:, 1].ravel()
    m[rz, 6] = c[:, 1].ravel()
    m[rz, 7] = s[:, 1].ravel()
    return ops


def request_coeffs(req: Request) -> np.ndarray:
    """Coefficients of request_terms(req) (the terms depend on n and kind only)."""
    n, g = req.n, req.g
    if req.ham == "tfim":
        return np.concatenate([np.full(n - 1, -1.0), np.full(n, -g)])
    return np.concatenate([np.ones(3 * n), np.full(n, g)])


# Requests of at most this many qubits run through the batched launch
# (tqsv_batch_adjoint): one work-group per request, state and lambda in LDS.
BATCH_MAX_QUBITS = 13
BATCH_MAX_REQUESTS = 512


def is_small(req: Request) -> bool:
    return req.n <= BATCH_MAX_QUBITS


class Cache:
    """Resources reused across requests: plans per (n, depth) (a new request
    only re-emits the programs) and Hamiltonians per (n, kind) (a new request
    only replaces the coefficients).  Several requests of one shape can be in
    flight together (a batch), so each key holds a pool."""

    def __init__(self):
        self.plans: Dict[Tuple[int, int], List["nat.Plan"]] = {}
        self.hams: Dict[Tuple[int, str], List["nat.Ham"]] = {}
        self.batch: Optional["nat.Batch"] = None

    def acquire(self, req: Request):
        ops = request_ops(req)
        pool = self.plans.setdefault((req.n, req.depth), [])
        if pool:
            plan = pool.pop()
            plan.update(ops)
        else:
            plan = nat.Plan(req.n, ops, jit=-1, min_qubits=BATCH_MAX_QUBITS if is_small(req) else 0)
        hpool = self.hams.setdefault((req.n, req.ham), [])
        if hpool:
            ham = hpool.pop()
            ham.set_coeffs(request_coeffs(req))
        else:
            x, z, c = PauliSum(req.n, request_terms(req)).to_masks()
            ham = nat.Ham(req.n, x, z, c)
        return plan, ham

    def release(self, req: Request, plan, ham) -> None:
        self.plans[(req.n, req.depth)].append(plan)
        self.hams[(req.n, req.ham)].append(ham)

    def get_batch(self) -> "nat.Batch":
        if self.batch is None:
            self.batch = nat.Batch()
        return self.batch


def execute_batch(reqs: List[Request], cache: Cache) -> List[Tuple[float, np.ndarray]]:
    """Energies and gradients of small requests in one launch (current stream)."""
    got = [cache.acquire(r) for r in reqs]
    try:
        psis = [nat.new_state(r.n, 0) for r in reqs]
        out = cache.get_batch().adjoint([p for p, _ in got], [h for _, h in got], psis,
                                        [len(r.params) for r in reqs])
    finally:
        for r, (p, h) in zip(reqs, got):
            cache.release(r, p, h)
    return [(float(e), g) for e, g in out]


def execute(req: Request, cache: Cache) -> Tuple[float, np.ndarray]:
    """Energy and gradient of one request on the current stream.  Small
    requests take the batched launch (a batch of one), so that a request's
    result does not depend on whether it was served alone or batched."""
    import torch

    if is_small(req):
        return execute_batch([req], cache)[0]
    plan, ham = cache.acquire(req)
    try:
        psi = nat.new_state(req.n, 0)
        lam = torch.empty_like(psi)
        e, g = plan.adjoint(ham, psi, lam, len(req.params))
    finally:
        cache.release(req, plan, ham)
    return float(e), g


def gpu_share(n: int) -> float:
    """Fraction of the chip the tile passes of an n-qubit request fill."""
    return min(1.0, float(2 ** max(int(n) - 13, 0)) / 512.0)


class Server:
    """Continuous-batching server: up to `slots` requests in flight, admitted
    in arrival order while their footprint fits `mem_budget` bytes and their
    GPU shares sum to at most `share_cap`.  `execute_fn(req, cache)` and
    `execute_batch_fn(reqs, cache)` default to the GPU paths; `streams=False`
    runs without HIP streams (host tests with stand-in executors)."""

    def __init__(self, slots: int = 8, mem_budget: float = 64e9, execute_fn=None,
                 execute_batch_fn=None, streams: bool = True, share_cap: float = 2.0):
        self.slots = int(slots)
        self.share_cap = float(share_cap)
        self.mem_budget = float(mem_budget)
        self._local = threading.local()
        self._execute = execute_fn or execute
        self._execute_batch = execute_batch_fn or execute_batch
        self._streams = bool(streams)
        self.last_stats: Dict[str, int] = {}

    @staticmethod
    def footprint(req: Request) -> float:
        return float(1 << req.n) * 24.0

    def _stream_ctx(self, stream):
        if stream is None:
            import contextlib

            return contextlib.nullcontext()
        import torch

        return torch.cuda.stream(stream)

    def _new_stream(self):
        if not self._streams:
            return None
        import torch

        return torch.cuda.Stream()

    def _run_one(self, req: Request) -> Tuple[float, float, np.ndarray]:
        w = getattr(self._local, "w", None)
        if w is None:
            w = (self._new_stream(), Cache())
            self._local.w = w
        start = time.perf_counter()
        with self._stream_ctx(w[0]):
            e, g = self._execute(req, w[1])
        return start, e, g

    def run(self, reqs: List[Request], realtime: bool = True) -> List[Result]:
        """Serve `reqs` (arrival times relative to the call) and return results
        in request order.  Requests of at most BATCH_MAX_QUBITS qubits go to
        the batch lane: whenever it is free, it takes every small request that
        has arrived (up to BATCH_MAX_REQUESTS) and serves them in one launch on
        its own stream, while larger requests run on the slot workers.  A
        failed request has `error` set in its result; a batch that fails is
        retried request by request, so one bad request fails alone."""
        pending = sorted(reqs, key=lambda r: r.arrival)
        results: Dict[int, Result] = {}
        t0 = time.perf_counter()
        lock = threading.Condition()
        state = {"mem": 0.0, "busy": 0, "share": 0.0}
        small_q: List[Tuple[Request, float]] = []
        lane = {"closed": False}
        self.last_stats = {"batches": 0, "batched_requests": 0, "rejected": 0, "failed": 0}

        def fail(req: Request, t_arr: float, start: float, exc: BaseException) -> Result:
            return Result(req.rid, float("nan"), np.zeros(0), t_arr, start, time.perf_counter(),
                          error=f"{type(exc).__name__}: {exc}")

        def lane_loop():
            stream = self._new_stream()
            cache = Cache()
            while True:
                with lock:
                    while not small_q and not lane["closed"]:
                        lock.wait()
                    if not small_q:
                        return
                    batch = small_q[:BATCH_MAX_REQUESTS]
                    del small_q[:len(batch)]
                start = time.perf_counter()
                self.last_stats["batches"] += 1
                self.last_stats["batched_requests"] += len(batch)
                done_list: List[Result] = []
                try:
                    with self._stream_ctx(stream):
                        out = self._execute_batch([r for r, _ in batch], cache)
                    done = time.perf_counter()
                    done_list = [Result(r.rid, e, g, ta, start, done)
                                 for (r, ta), (e, g) in zip(batch, out)]
                except Exception:
                    # isolate the failure: every request of the batch alone
                    for r, ta in batch:
                        try:
                            with self._stream_ctx(stream):
                                ((e, g),) = self._execute_batch([r], cache)
                            done_list.append(Result(r.rid, e, g, ta, start, time.perf_counter()))
                        except Exception as exc:
                            done_list.append(fail(r, ta, start, exc))
                with lock:
                    for res in done_list:
                        results[res.rid] = res
                        self.last_stats["failed"] += 0 if res.ok else 1
                    lock.notify_all()

        lane_thread = threading.Thread(target=lane_loop, daemon=True)
        lane_thread.start()

        def done_cb(req, fut, t_arr):
            try:
                start, e, g = fut.result()
                res = Result(req.rid, e, g, t_arr, start, time.perf_counter())
            except Exception as exc:
                res = fail(req, t_arr, t_arr, exc)
            with lock:
                results[req.rid] = res
                self.last_stats["failed"] += 0 if res.ok else 1
                state["mem"] -= self.footprint(req)
                state["share"] -= gpu_share(req.n)
                state["busy"] -= 1
                lock.notify_all()

        with ThreadPoolExecutor(max_workers=self.slots) as ex:
            i = 0
            while i < len(pending):
                req = pending[i]
                t_arr = t0 + req.arrival if realtime else t0
                now = time.perf_counter()
                if realtime and now < t_arr:
                    time.sleep(min(t_arr - now, 0.002))
                    continue
                i += 1
                if self.footprint(req) > self.mem_budget:
                    with lock:
                        results[req.rid] = Result(
                            req.rid, float("nan"), np.zeros(0), t_arr, t_arr, t_arr,
                            error=f"rejected: adjoint footprint {self.footprint(req):.3g} B "
                                  f"exceeds the memory budget {self.mem_budget:.3g} B")
                        self.last_stats["rejected"] += 1
                    continue
                if is_small(req):
                    with lock:
                        small_q.append((req, t_arr))
                        lock.notify_all()
                    continue
                share = gpu_share(req.n)
                with lock:
                    while state["busy"] >= self.slots or (
                            state["busy"] > 0 and (
                                state["mem"] + self.footprint(req) > self.mem_budget
                                or state["share"] + share > self.share_cap + 1e-9)):
                        lock.wait()
                    state["busy"] += 1
                    state["mem"] += self.footprint(req)
                    state["share"] += share
                fut = ex.submit(self._run_one, req)
                fut.add_done_callback(lambda f, r=req, ta=t_arr: done_cb(r, f, ta))
        with lock:
            lane["closed"] = True
            lock.notify_all()
        lane_thread.join()
        return [results[r.rid] for r in reqs]


def plain_loop(reqs: List[Request], realtime: bool = True, execute_fn=None,
               sync: bool = True) -> List[Result]:
    """The same requests one at a time, in arrival order, on one stream."""
    run = execute_fn or execute
    cache = Cache()
    out: Dict[int, Result] = {}
    t0 = time.perf_counter()
    for req in sorted(reqs, key=lambda r: r.arrival):
        t_arr = t0 + req.arrival if realtime else t0
        now = time.perf_counter()
        if realtime and now < t_arr:
            time.sleep(t_arr - now)
        start = time.perf_counter()
        try:
            e, g = run(req, cache)
            if sync:
                import torch

                torch.cuda.current_stream().synchronize()
            out[req.rid] = Result(req.rid, e, g, t_arr, start, time.perf_counter())
        except Exception as exc:
            out[req.rid] = Result(req.rid, float("nan"), np.zeros(0), t_arr, start,
                                  time.perf_counter(), error=f"{type(exc).__name__}: {exc}")
    return [out[r.rid] for r in reqs]


def summarize(results: List[Result]) -> Dict[str, float]:
    """Throughput and latency percentiles over the requests that succeeded."""
    good = [r for r in results if r.ok]
    out: Dict[str, float] = {"requests": len(results), "failed": len(results) - len(good)}
    if not good:
        return out
    lat = np.array([r.latency for r in good])
    span = max(r.done for r in good) - min(r.arrival for r in good)
    out.update({"throughput_rps": round(len(good) / span, 2),
                "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)})
    return out
