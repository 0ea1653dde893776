#!/usr/bin/env python3
"""Benchmark of the MI355X statevector engine (one GPU).

Default workload (--workload brickwork): the 30-qubit complex64 brickwork of
20 layers (RX.RZ on every qubit, CZ on even/odd pairs) from |0...0>, then the
norm and the energy of a random 500-term 2-local Pauli sum, both reduced in
FP64.  One step = that whole job with the state resident in HBM.

`value` (GB/s) = the step's layer-equivalent HBM traffic over the step time:
the initial state write (8 B/amp), one read and one write of the state per
gate layer (16 B/amp per layer, what a layer-by-layer engine has to move) and
one read for the expectation (8 B/amp).  Fused passes move less than that, so
value / 8000 is the "fraction of HBM peak per gate layer" of the target.
`roofline` is the measured per-launch bandwidth of the dominant kernel (the
fused gate pass: one read and one write of the state per launch) from HIP
events on its launch stream; `traffic` comes from the committed rocprofv3
FETCH_SIZE / WRITE_SIZE summary when one exists for this workload.
`cpu_baseline` times the oracle (NumPy complex128, the reference's per-gate
contraction) on a bounded sample of the same workload on the host cores.

The default line also carries the other two clauses of the metric as extra
objects: `adjoint_22q` (the 22-qubit adjoint energy+gradient rate against the
plain per-parameter GPU path, its error against the stored complex128 result
and the dual pass's roofline) and `serve` (the continuous-batching server
against the one-at-a-time loop at one offered rate); --no-extras leaves them
out.

Other workloads: --workload adjoint (22-qubit RY.RZ + CX-ring ansatz, 528
parameters, periodic Heisenberg chain), hea (12-qubit hardware-efficient
ansatz, open-chain TFIM energy+gradient), ucc (UCCSD, 20 spin-orbitals, on
generated kernels), serve (per-rate table of the server against the plain
loop), sample (10^6 shots from a 30-qubit state), mfma (a dense 5-qubit block
through the vendor FP32 matrix-core GEMM against the fused VALU pass).

    python bench.py [--steps K] [--warmup W] [--workload brickwork|adjoint|hea|ucc|serve|sample|mfma]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBPS = 8000.0
METRIC = "30q gate GB/s (% HBM peak); VQE energy+grad/s and p99 latency vs plain"
ROOT = os.path.dirname(os.path.abspath(__file__))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="brickwork",
                    choices=["brickwork", "adjoint", "hea", "ucc", "serve", "sample", "mfma"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--layers", type=int, default=None)
    ap.add_argument("--terms", type=int, default=500)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--tile-bits", type=int, default=0)
    ap.add_argument("--max-cost", type=int, default=None,
                    help="per-pass VALU budget (ops per amplitude); default 260 for the 30q "
                         "brickwork, the planner's default elsewhere")
    ap.add_argument("--jit", type=int, default=0, help="1 generated kernels, -1 precompiled, 0 auto")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed steps, copy the state once with a device copy "
                         "kernel of known size (calibration of HBM byte counters)")
    ap.add_argument("--rates", default="50,100,200,500,1000,2000",
                    help="serve workload: offered request rates (req/s)")
    ap.add_argument("--requests", type=int, default=120, help="serve workload: requests per rate")
    ap.add_argument("--slots", type=int, default=8, help="serve workload: requests in flight")
    ap.add_argument("--shots", type=int, default=1_000_000, help="sample workload: shots per call")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload: leave out the adjoint and serving parts of the line")
    ap.add_argument("--serve-rate", type=float, default=200.0,
                    help="default workload: offered rate of the serving part (req/s)")
    ap.add_argument("--shift-sample", type=int, default=8,
                    help="adjoint workload: parameters timed on the parameter-shift path")
    return ap.parse_args()


def _sync():
    import torch

    torch.cuda.synchronize()


def _events():
    import torch

    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _threads() -> int:
    try:
        from threadpoolctl import threadpool_info

        return int(max([d.get("num_threads", 1) for d in threadpool_info()] or [1]))
    except Exception:
        return 1


def _golden(name: str):
    """Complex128 results of the same workload (tests/golden/fullsize_c128.json)."""
    try:
        with open(os.path.join(ROOT, "tests", "golden", "fullsize_c128.json")) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


# ---- config 2: 30-qubit brickwork + 500-term <H> ----------------------------------
def cpu_baseline_brickwork(n: int, layers: int, nterms: int) -> dict:
    """The oracle (NumPy complex128, the reference's per-gate contraction and
    per-term Pauli application) on a bounded sample of the workload on the
    host cores: RX (3 runs, median), RZ and CZ on the middle qubits and one
    2-local term <X_i Y_j>, each on the full 2^n state.  A layer of the
    reference's dispatch is n RX + n RZ + about n/2 CZ; the job is `layers`
    layers plus `nterms` terms.  BLAS threads = all host cores."""
    import contextlib

    from oracle import statevector as O

    try:
        from threadpoolctl import threadpool_limits

        ctx = threadpool_limits(limits=os.cpu_count() or 1)
    except Exception:
        ctx = contextlib.nullcontext()
    with ctx:
        psi = np.zeros(1 << n, dtype=np.complex128)
        psi[0] = 1.0
        q = n // 2
        t_rx = []
        for k in range(3):
            t0 = time.perf_counter()
            psi = O.apply_1q(psi, O.gate_rx(0.3 + k), q, n)
            t_rx.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        psi = O.apply_1q(psi, O.gate_rz(0.7), q, n)
        t_rz = time.perf_counter() - t0
        t0 = time.perf_counter()
        psi = O.apply_2q(psi, O.gate_cz(), q, q + 1, n)
        t_cz = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.pauli_sum_expectation(psi, [(1.0, [("X", q - 1), ("Y", q + 2)])], n)
        t_term = time.perf_counter() - t0
        del psi
        threads = _threads()
    t1 = float(np.median(t_rx))
    layer_s = n * t1 + n * t_rz + (n // 2) * t_cz
    job_s = layers * layer_s + nterms * t_term
    step_bytes = 8 * (1 << n) + layers * 16 * (1 << n) + 8 * (1 << n)
    return {
        "value": round(step_bytes / job_s / 1e9, 4),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "cpu": _cpu_model(),
        "sample": (f"extrapolated: NumPy complex128 per-gate contraction and per-term Pauli "
                   f"application (the oracle's restatement of apply_1q/apply_2q_statevector and "
                   f"the Pauli-sum expectation) timed on the full {n}-qubit state: RX q{q} "
                   f"{t1:.2f} s (median of 3), RZ {t_rz:.2f} s, CZ q{q},{q + 1} {t_cz:.2f} s, one "
                   f"2-local term {t_term:.2f} s; job = {layers} layers x ({n} RX + {n} RZ + "
                   f"{n // 2} CZ) + {nterms} terms = {job_s:.0f} s; value counts the same "
                   f"layer-equivalent bytes as the GPU value"),
        "job_s": round(job_s, 1),
    }


def run_brickwork(args) -> dict:
    import torch

    from tyxonq_amd import _native as nat
    from tyxonq_amd import bench_workloads as W
    from tyxonq_amd.engine import lower
    from tyxonq_amd.pauli import PauliSum

    n = args.n or 30
    layers = args.layers or 20
    circ = W.brickwork_circuit(n, layers, args.seed)
    terms = W.random_2local_terms(n, args.terms, args.seed)
    L, _, _ = lower(circ)
    ops = L.segments[0][1]
    t0 = time.perf_counter()
    if args.max_cost is None:
        args.max_cost = 260
    plan = nat.Plan(n, ops, tile_bits=args.tile_bits, max_cost=args.max_cost, jit=args.jit)
    x, z, c = PauliSum(n, terms).to_masks()
    ham = nat.Ham(n, x, z, c)
    build_s = time.perf_counter() - t0
    info = plan.info()
    psi = torch.empty(1 << n, dtype=torch.complex64, device=nat.device())

    for _ in range(args.warmup):
        nat.init_basis(psi, 0)
        plan.run(psi)
        ham.expval(psi)
    _sync()
    gate_ms, exp_ms, energies = [], [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        nat.init_basis(psi, 0)
        e0, e1 = _events()
        e0.record()
        plan.run(psi)
        e1.record()
        e2, e3 = _events()
        e2.record()
        energies.append(ham.expval(psi))  # synchronizes the stream
        e3.record()
        e3.synchronize()
        exp_ms.append(e2.elapsed_time(e3))
        gate_ms.append(e0.elapsed_time(e1))
    _sync()
    ms_per_step = (time.perf_counter() - t_start) / args.steps * 1e3
    dim = 1 << n
    passes = info["passes"]
    step_bytes = 8 * dim + layers * 16 * dim + 8 * dim
    value = step_bytes / (ms_per_step / 1e3) / 1e9
    avg_gate_ms = float(np.mean(gate_ms))
    per_launch_ms = avg_gate_ms / passes
    achieved = 16 * dim / (per_launch_ms / 1e3) / 1e9
    energy, norm = energies[-1]
    workload = f"brickwork {n}q x {layers} layers + {args.terms}-term 2-local <H>"
    gold = _golden("brickwork30")
    vs_c128 = None
    if gold and (gold["qubits"], gold["layers"], gold["seed"], gold["pauli_terms"]) == \
            (n, layers, args.seed, args.terms):
        vs_c128 = {"abs_dE": abs(energy - gold["energy"]), "abs_dnorm2": abs(norm - gold["norm2"]),
                   "energy_c128": gold["energy"]}
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded brickwork circuit and random 2-local Pauli sum)",
        "config": {"workload": workload, "qubits": n, "layers": layers,
                   "gates": int(len(ops)), "pauli_terms": args.terms, "seed": args.seed,
                   "reductions": "fp64", "parallelism": "1 GPU"},
        "value_frac_of_peak": round(value / HBM_PEAK_GBPS, 4),
        "gate_layer_GBps": round(layers * 16 * dim / (avg_gate_ms / 1e3) / 1e9, 1),
        "gate_ms": round(avg_gate_ms, 3),
        "layer_ms": round(avg_gate_ms / layers, 3),
        "expval_ms": round(float(np.mean(exp_ms)), 3),
        "gate_passes": passes,
        "expval_passes": int(ham.info()["passes"]),
        "plan": {"tile_bits": info["tile_bits"], "phases": info["phases"],
                 "gates_after_merge": info["gates_after_merge"], "max_cost": args.max_cost,
                 "generated_kernels": args.jit >= 0, "build_s": round(build_s, 1)},
        "energy": energy,
        "norm2": norm,
        "vs_complex128": vs_c128,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": _traffic(f"gate_pass_{n}q"),
            "kernel": "tqsv_gate_pass (fused tile pass, one state read + write)",
            "bytes_per_launch": 16 * dim,
            "avg_launch_ms": round(per_launch_ms, 4),
        },
    }
    if args.calib_copy:
        other = torch.empty_like(psi)
        other.copy_(psi)
        _sync()
        del other
    del psi
    torch.cuda.empty_cache()
    return out


# ---- config 3: adjoint energy+gradient vs the plain per-parameter path --------------
def run_adjoint(args, steps=None, warmup=None, n=None, layers=None) -> dict:
    import torch

    from tyxonq_amd import _native as nat
    from tyxonq_amd import bench_workloads as W
    from tyxonq_amd.engine import lower
    from tyxonq_amd.pauli import PauliSum

    n = n or args.n or 22
    layers = layers or args.layers or 12
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    tt = torch.tensor(np.random.default_rng(args.seed).normal(0.0, 0.5, size=2 * n * layers),
                      dtype=torch.float64, requires_grad=True)
    circ, _ = W.ryrz_ring_ansatz(n, layers, params=[tt[i] for i in range(tt.numel())])
    terms = W.heisenberg_ring_terms(n, 0.5)
    L, _, _ = lower(circ, track_grad=True)
    ops = L.segments[0][1]
    t0 = time.perf_counter()
    plan = nat.Plan(n, ops, tile_bits=args.tile_bits or 0, max_cost=args.max_cost or 0, jit=args.jit)
    x, z, c = PauliSum(n, terms).to_masks()
    ham = nat.Ham(n, x, z, c)
    dev = nat.device()
    psi = torch.empty(1 << n, dtype=torch.complex64, device=dev)
    lam = torch.empty_like(psi)
    nat.init_basis(psi, 0)
    plan.adjoint(ham, psi, lam)  # builds the reverse plan
    build_s = time.perf_counter() - t0
    npar = plan.nparams
    for _ in range(warmup):
        nat.init_basis(psi, 0)
        plan.adjoint(ham, psi, lam)
    _sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        nat.init_basis(psi, 0)
        e, g = plan.adjoint(ham, psi, lam)
    _sync()
    adj_s = (time.perf_counter() - t0) / steps
    roof = _dual_pass_roofline(plan, psi, lam, n)
    gold = _golden("adjoint22")
    vs_c128 = None
    if gold and (gold["qubits"], gold["layers"], gold["seed"]) == (n, layers, args.seed):
        vs_c128 = {"abs_dE": abs(e - gold["energy"]),
                   "max_abs_dgrad": float(np.max(np.abs(g - np.asarray(gold["grad"])))),
                   "params_compared": len(gold["grad"])}

    # plain path: parameter shift with the same kernels, 2 energies per parameter
    base = ops.copy()
    k = min(args.shift_sample, npar)
    splan = nat.Plan(n, ops, max_cost=args.max_cost or 0, jit=args.jit)
    sel = np.nonzero(base["param"] >= 0)[0]

    th = tt.detach().numpy()

    def energy_with(row, delta):
        """Energy with the parameter of op `row` shifted by delta (RY or RZ)."""
        o = base.copy()
        ang = th[int(o[row]["param"])] + delta
        cs, sn = np.cos(ang / 2), np.sin(ang / 2)
        if o[row]["gx"] != 0:  # RY: generator Y
            m = [cs, 0.0, -sn, 0.0, sn, 0.0, cs, 0.0]
        else:  # RZ: generator Z
            m = [cs, -sn, 0.0, 0.0, 0.0, 0.0, cs, sn]
        o[row]["m"][:8] = m
        ta = time.perf_counter()
        splan.update(o)
        _sync()
        tb = time.perf_counter()
        nat.init_basis(psi, 0)
        splan.run(psi)
        _sync()
        tc = time.perf_counter()
        en = ham.expval(psi)[0]
        td = time.perf_counter()
        parts["update"] += tb - ta
        parts["run"] += tc - tb
        parts["expval"] += td - tc
        return en

    parts = {"update": 0.0, "run": 0.0, "expval": 0.0}

    grad_ps = np.zeros(k)
    energy_with(sel[0], 0.0)  # warm the update path
    _sync()
    parts = {"update": 0.0, "run": 0.0, "expval": 0.0}
    t0 = time.perf_counter()
    for j in range(k):
        grad_ps[j] = 0.5 * (energy_with(sel[j], np.pi / 2) - energy_with(sel[j], -np.pi / 2))
    _sync()
    ps_per_param = (time.perf_counter() - t0) / k
    plain_s = ps_per_param * npar + ps_per_param / 2  # 2P shifted energies + one energy
    # the same with the host's plan updates left out: state preparation and
    # expectation only (the GPU work of the plain path)
    gpu_per_param = (parts["run"] + parts["expval"]) / k
    plain_gpu_s = gpu_per_param * npar + gpu_per_param / 2
    return {
        "metric": METRIC,
        "value": round(1.0 / adj_s, 3),
        "unit": "energy+grad/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(adj_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded RY.RZ + CX-ring ansatz, periodic Heisenberg chain, hz=0.5)",
        "config": {"workload": f"adjoint {n}q x {layers} layers, {npar} params, "
                               f"{len(terms)}-term Heisenberg ring", "qubits": n,
                   "layers": layers, "params": npar, "reductions": "fp64",
                   "parallelism": "1 GPU"},
        "energy": e,
        "plain_path": {"kind": "parameter shift on the same kernels (2 energies per parameter)",
                       "value": round(1.0 / plain_s, 5), "unit": "energy+grad/s",
                       "timed_params": k, "s_per_param": round(ps_per_param, 4),
                       "ms_per_param": {kk: round(v / k * 1e3, 3) for kk, v in parts.items()},
                       "gpu_only_value": round(1.0 / plain_gpu_s, 5)},
        "speedup_vs_plain": round(plain_s / adj_s, 1),
        "speedup_vs_plain_gpu_only": round(plain_gpu_s / adj_s, 1),
        "max_abs_grad_diff_vs_shift_sample": float(np.max(np.abs(g[:k] - grad_ps))) if k else None,
        "vs_complex128": vs_c128,
        "build_s": round(build_s, 1),
        "roofline": roof,
        "cpu_baseline": None,
    }


def _dual_pass_roofline(plan, psi, lam, n: int) -> dict:
    """The dominant kernel of an energy+gradient: the dual pass of the reverse
    sweep (psi and lambda read and written once per launch, 32 bytes per
    amplitude), timed alone through the VJP entry, which runs exactly the
    reverse sweep; the average over its passes, per-pass partial sums
    included."""
    info = plan.info()
    nrev = max(int(info["reverse_passes"]), 1)
    ev0, ev1 = _events()
    vjp_ms = []
    for _ in range(5):
        ev0.record()
        plan.vjp(psi, lam, plan.nparams)
        ev1.record()
        ev1.synchronize()
        vjp_ms.append(ev0.elapsed_time(ev1))
    dual_ms = float(np.median(vjp_ms)) / nrev
    dual_bytes = 32 * (1 << n)
    achieved = dual_bytes / (dual_ms * 1e-3) / 1e9
    gen = bool(info.get("reverse_generated", False))
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
            "kernel": ("tqsv_adjoint_pass" if gen else "k_pass<8, 2, ...>")
            + " (dual pass: psi and lambda read + written)",
            "bytes_per_launch": dual_bytes, "avg_launch_ms": round(dual_ms, 5),
            "launches_per_sweep": nrev}


# ---- config 1: 12-qubit hardware-efficient ansatz, TFIM --------------------------
def run_hea(args) -> dict:
    import torch

    import tyxonq_amd as tq
    from tyxonq_amd import bench_workloads as W
    from tyxonq_amd.engine import StatevectorEngine

    n = args.n or 12
    layers = args.layers or 4
    th0 = np.random.default_rng(args.seed).uniform(0.0, 2.0 * np.pi, size=n * layers)
    terms = W.tfim_terms(n, 1.0)
    eng = StatevectorEngine(backend_name="pytorch")

    def step():
        tt = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
        c, _ = W.hea_ry_cx(n, layers, params=[tt[i] for i in range(tt.numel())])
        return eng.energy_and_grad(c, terms)

    for _ in range(args.warmup):
        step()
    _sync()
    lat = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        e, g = step()
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat)
    return {
        "metric": METRIC,
        "value": round(1.0 / float(lat.mean()), 2),
        "unit": "energy+grad/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(lat.mean()) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded RY + CX-ladder ansatz, open-chain TFIM g=1)",
        "config": {"workload": f"hea {n}q x {layers} layers, TFIM energy+grad", "qubits": n,
                   "layers": layers, "params": n * layers, "parallelism": "1 GPU"},
        "energy": e,
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "roofline": None,
        "cpu_baseline": None,
    }


# ---- config 4: UCCSD, 20 spin orbitals ----------------------------------------
def run_ucc(args) -> dict:
    """One step = new parameters -> energy and dE/dtheta for the 450 shared
    parameters: the parameter update re-emits the plan's programs (selection
    kept), then one forward and one adjoint sweep.  The same step through the
    Circuit / torch.autograd surface is timed as `autograd_path`."""
    import torch

    from tyxonq_amd import _native as nat
    from tyxonq_amd import chem
    from tyxonq_amd.engine import StatevectorEngine
    from tyxonq_amd.pauli import PauliSum

    ns = (args.n or 20) // 2
    no = ns // 2
    nv = ns - no
    n = 2 * ns
    t0 = time.perf_counter()
    h, g = chem.synthetic_integrals(ns, args.seed)
    terms = chem.molecular_hamiltonian(h, g)
    ex, ids = chem.uccsd_excitations(no, nv)
    ans = chem.UCCAnsatz(n, (no, no), ex, ids)
    npar = ans.nparams
    rng = np.random.default_rng(args.seed)
    thetas = [rng.normal(0.0, 0.05, size=npar) for _ in range(args.warmup + args.steps)]
    x, z, c = PauliSum(n, terms).to_masks()
    ham = nat.Ham(n, x, z, c)
    # generated kernels unless --jit -1: about 160 pass kernels, compiled once
    # per plan structure (seconds on 16 host threads), then reused for every
    # parameter update
    jit = args.jit if args.jit else 1
    plan = nat.Plan(n, ans.ops(thetas[0]), tile_bits=args.tile_bits or 13, max_cost=100000,
                    jit=jit)
    psi = torch.empty(1 << n, dtype=torch.complex64, device=nat.device())
    lam = torch.empty_like(psi)
    nat.init_basis(psi, 0)
    e0, g0 = plan.adjoint(ham, psi, lam)
    build_s = time.perf_counter() - t0
    gold = _golden("ucc20")
    vs_c128 = None
    if gold and (gold["spin_orbitals"], gold["seed"], gold["theta_sigma"]) == (n, args.seed, 0.05):
        vs_c128 = {"abs_dE": abs(e0 - gold["energy"]),
                   "max_abs_dgrad": float(np.max(np.abs(g0 - np.asarray(gold["grad"])))),
                   "params_compared": len(gold["grad"])}

    opbuf = ans.ops(thetas[0])

    def step(th):
        plan.update(ans.ops(th, out=opbuf))
        nat.init_basis(psi, 0)
        return plan.adjoint(ham, psi, lam)

    for k in range(args.warmup):
        step(thetas[k])
    _sync()
    lat = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        e, gr = step(thetas[args.warmup + k])
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat)
    info = plan.info()
    roof = _dual_pass_roofline(plan, psi, lam, n)

    # the same energy+gradient through Circuit + torch.autograd (the drop-in surface)
    eng = StatevectorEngine(backend_name="pytorch")
    tt = torch.tensor(thetas[-1], dtype=torch.float64, requires_grad=True)
    circ = chem.ucc_circuit([tt[i] for i in range(npar)], n, (no, no), ex, ids)
    ea = eng.energy(circ, terms)
    ea.backward()
    t0 = time.perf_counter()
    tt = torch.tensor(thetas[-1], dtype=torch.float64, requires_grad=True)
    circ = chem.ucc_circuit([tt[i] for i in range(npar)], n, (no, no), ex, ids)
    ea = eng.energy(circ, terms)
    ea.backward()
    auto_s = time.perf_counter() - t0
    return {
        "metric": METRIC,
        "value": round(1.0 / float(lat.mean()), 3),
        "unit": "energy+grad/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(lat.mean()) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded integrals, HF reference, UCCSD theta ~ N(0, 0.05^2))",
        "config": {"workload": f"ucc {n} spin orbitals, {2 * no} electrons, {len(ex)} excitations "
                               f"({npar} params), {ans.nrot} Pauli rotations, {len(terms)}-term JW "
                               f"Hamiltonian", "qubits": n, "params": npar, "rotations": ans.nrot,
                   "pauli_terms": len(terms), "reductions": "fp64", "parallelism": "1 GPU"},
        "energy": e,
        "grad_norm": float(np.linalg.norm(gr)),
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "gate_passes": info["passes"],
        "generated_kernels": bool(info.get("generated", False)),
        "vs_complex128": vs_c128,
        "autograd_path": {"ms": round(auto_s * 1e3, 1),
                          "abs_dE_vs_direct": abs(float(ea.detach()) - e) if args.steps else None,
                          "max_abs_dgrad_vs_direct": float(np.max(np.abs(tt.grad.numpy() - gr)))},
        "build_s": round(build_s, 1),
        "roofline": roof,
        "cpu_baseline": None,
    }


# ---- config 5: serving with continuous batching ---------------------------------
def run_serve(args) -> dict:
    """Per offered rate: the same Poisson stream served by the continuous-
    batching server and by the plain one-request-at-a-time loop."""
    import torch

    from tyxonq_amd import serving as S

    rates = [float(r) for r in args.rates.split(",")]
    warm = S.make_requests(2 * args.slots, 1e6, seed=args.seed + 99)
    S.Server(slots=args.slots).run(warm, realtime=False)
    S.plain_loop(warm[:4], realtime=False)
    table = []
    for k, rate in enumerate(rates):
        reqs = S.make_requests(args.requests, rate, seed=args.seed + k)
        server = S.Server(slots=args.slots)
        srv = S.summarize(server.run(reqs))
        srv.update(server.last_stats)
        plain = S.summarize(S.plain_loop(reqs))
        table.append({"rate_rps": rate, "server": srv, "plain": plain})
        print(json.dumps({"rate": rate, "server": srv, "plain": plain}), file=sys.stderr, flush=True)
    top = table[-1]
    torch.cuda.synchronize()
    return {
        "metric": METRIC,
        "value": top["server"]["throughput_rps"],
        "unit": "energy+grad requests/s",
        "n_gpus": 1,
        "steps": args.requests,
        "warmup": len(warm),
        "ms_per_step": round(1e3 / top["server"]["throughput_rps"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded Poisson arrivals; n ~ U{10..26}, depth ~ U{2..16}, "
                "RY/RZ + CX-ladder ansatz, TFIM or Heisenberg with g ~ U[0.5, 1.5])",
        "config": {"workload": f"serving at {top['rate_rps']:.0f} req/s offered "
                               f"({args.requests} requests per rate), adjoint energy+grad",
                   "slots": args.slots, "admission": "2^n x 24 B footprint",
                   "batched": "requests of <= 13 qubits: one launch per batch, one "
                              "work-group per request",
                   "parallelism": "1 GPU, one HIP stream per slot and one for the batch lane"},
        "p50_ms": top["server"]["p50_ms"],
        "p99_ms": top["server"]["p99_ms"],
        "plain": top["plain"],
        "throughput_vs_plain": round(top["server"]["throughput_rps"] / top["plain"]["throughput_rps"], 2),
        "sweep": table,
        "roofline": None,
        "cpu_baseline": None,
    }


def serve_at(rate: float, requests: int, slots: int, seed: int) -> dict:
    """The serving workload at one offered rate: the continuous-batching server
    and the plain one-at-a-time loop on the same Poisson stream."""
    from tyxonq_amd import serving as S

    warm = S.make_requests(2 * slots, 1e6, seed=seed + 99)
    S.Server(slots=slots).run(warm, realtime=False)
    S.plain_loop(warm[:4], realtime=False)
    reqs = S.make_requests(requests, rate, seed=seed)
    server = S.Server(slots=slots)
    srv = S.summarize(server.run(reqs))
    srv.update(server.last_stats)
    plain = S.summarize(S.plain_loop(reqs))
    return {"rate_rps": rate, "requests": requests, "slots": slots, "server": srv,
            "plain": plain,
            "throughput_vs_plain": round(srv["throughput_rps"] / plain["throughput_rps"], 3),
            "p99_vs_plain": round(srv["p99_ms"] / plain["p99_ms"], 3)}


# ---- dense 5-qubit blocks: matrix cores against the VALU paths ---------------------
def run_mfma(args) -> dict:
    """Dense 5-qubit blocks on the 30-qubit state.  A 32x32 complex unitary
    (any number of gates on 5 qubits fuses into one; here a random one) is
    applied to the 5 lowest index bits of all 2^25 blocks (a) as one complex
    GEMM on the matrix cores (torch.matmul on complex64 -> the vendor BLAS
    cgemm, FP32 MFMA), (b) by the VALU k-qubit kernel of this library
    (tqsv_apply_kq, one output amplitude per thread), and both are set
    against (c) the fused tile pass of the 30-qubit brickwork plan, which
    applies about 28 single-qubit gates (RX.RZ) plus CZs per amplitude per
    pass.  A dense block costs 32 complex MACs (256 flops) per amplitude
    whatever it fuses, and covers only the 5 qubits of its block."""
    import torch

    from tyxonq_amd import _native as nat
    from tyxonq_amd import bench_workloads as W
    from tyxonq_amd.engine import lower

    n = args.n or 30
    dim = 1 << n
    dev = nat.device()
    rng = np.random.default_rng(args.seed)
    A = rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32))
    U = np.linalg.qr(A)[0]
    g = torch.Generator(device=dev).manual_seed(args.seed)
    psi = torch.complex(torch.randn(dim, device=dev, generator=g),
                        torch.randn(dim, device=dev, generator=g)).to(torch.complex64)
    psi /= torch.linalg.vector_norm(psi)
    UT = torch.tensor(U.T.copy(), dtype=torch.complex64, device=dev)
    out = torch.empty_like(psi)
    ev0, ev1 = _events()
    reps = max(args.steps, 3)

    def timed(fn):
        for _ in range(max(args.warmup, 1)):
            fn()
        _sync()
        ev0.record()
        for _ in range(reps):
            fn()
        ev1.record()
        ev1.synchronize()
        return ev0.elapsed_time(ev1) / reps

    ms_mfma = timed(lambda: torch.matmul(psi.view(dim // 32, 32), UT, out=out.view(dim // 32, 32)))
    got_mfma = out[: 1 << 20].clone()
    res = {}

    def valu():
        res["o"] = nat.apply_kq(psi, [4, 3, 2, 1, 0], U)

    ms_valu = timed(valu)
    diff = float((res["o"][: 1 << 20] - got_mfma).abs().max().item())
    del res
    torch.cuda.empty_cache()
    circ = W.brickwork_circuit(n, 20, args.seed)
    L, _, _ = lower(circ)
    plan = nat.Plan(n, L.segments[0][1], max_cost=260)
    passes = int(plan.info()["passes"])
    ms_plan = timed(lambda: plan.run(psi)) if passes else 0.0
    nbytes = 16 * dim
    flops = 256 * dim

    def line(ms):
        return {"ms": round(ms, 4), "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1),
                "TFLOPs": round(flops / (ms * 1e-3) / 1e12, 2)}

    a, b = line(ms_mfma), line(ms_valu)
    per_pass = ms_plan / max(passes, 1)
    return {
        "metric": METRIC,
        "value": a["GBps"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": args.warmup,
        "ms_per_step": a["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded random state and random 32x32 unitary)",
        "config": {"workload": f"dense 32x32 complex block on the 5 low bits of a {n}-qubit state",
                   "qubits": n, "parallelism": "1 GPU"},
        "mfma_cgemm": a,
        "valu_kq_kernel": b,
        "max_abs_diff_mfma_vs_valu": diff,
        "fused_pass": {"ms": round(per_pass, 4),
                       "GBps": round(nbytes / (per_pass * 1e-3) / 1e9, 1),
                       "single_qubit_gates_per_pass": round(2 * n * 20 / max(passes, 1), 1),
                       "passes": passes, "plan_ms": round(ms_plan, 3)},
        "roofline": {"bound": "hbm", "achieved": a["GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(a["GBps"] / HBM_PEAK_GBPS, 4), "traffic": None,
                     "kernel": "vendor cgemm (FP32 MFMA) applying the dense block",
                     "bytes_per_launch": nbytes, "avg_launch_ms": a["ms"]},
        "cpu_baseline": None,
    }


# ---- shot sampling (counts path of engine.run) ------------------------------------
def run_sample(args) -> dict:
    """Counts of `shots` draws from a 30-qubit state (8 GiB, random amplitudes)
    with the GPU sampler; the state is resident before the timed calls."""
    import torch

    from tyxonq_amd import _native as nat

    n = args.n or 30
    shots = int(args.shots)
    dev = nat.device()
    g = torch.Generator(device=dev).manual_seed(args.seed)
    psi = torch.complex(torch.randn(1 << n, device=dev, generator=g),
                        torch.randn(1 << n, device=dev, generator=g)).to(torch.complex64)
    smp = nat.Sampler(n, shots)
    for w in range(args.warmup):
        smp.sample(psi, shots, seed=w)
    _sync()
    s0, s1 = _events()
    s0.record()
    t0 = time.perf_counter()
    for k in range(args.steps):
        idx, cnt = smp.sample(psi, shots, seed=1000 + k)
    s1.record()
    _sync()
    wall = (time.perf_counter() - t0) / args.steps
    gpu_ms = s0.elapsed_time(s1) / args.steps
    assert int(cnt.sum()) == shots
    # algorithmic bytes: one read of the state (the segment and block sums);
    # the shots then read only the 64-amplitude segments they land in
    nbytes = 8 * (1 << n)
    achieved = nbytes / (gpu_ms * 1e-3) / 1e9
    return {
        "metric": METRIC,
        "value": round(shots / wall, 1),
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random complex64 amplitudes)",
        "config": {"workload": f"sample {shots} shots from a {n}-qubit state", "qubits": n,
                   "shots": shots, "unique_states": int(idx.size), "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(achieved / 8000.0, 4), "traffic": None,
                     "kernel": "whole sampler call (segment sums, scans, resolve, compaction, "
                               "copy of the counts to the host)",
                     "bytes_per_launch": nbytes, "avg_launch_ms": round(gpu_ms, 4)},
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_sample(n, shots),
    }


def cpu_baseline_sample(n: int, shots: int) -> dict:
    """The reference's host path (|psi|^2, choice, bincount) on a 26-qubit
    state, scaled by 2^(n-26) (the work is linear in the state size)."""
    from oracle import sampling as OS

    m = min(n, 26)
    rng = np.random.default_rng(0)
    v = (rng.normal(size=1 << m) + 1j * rng.normal(size=1 << m)).astype(np.complex64)
    t0 = time.perf_counter()
    OS.reference_counts(v, shots, seed=1)
    dt = (time.perf_counter() - t0) * (1 << (n - m))
    return {"value": round(shots / dt, 1), "unit": "shots/s", "cores": 1, "kind": "port",
            "sample": f"numpy |psi|^2 + Generator.choice + bincount (engine.py:377-465) on a "
                      f"{m}-qubit state with {shots} shots: {dt / (1 << (n - m)):.2f} s, "
                      f"scaled x{1 << (n - m)} to {n} qubits"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                                   "(how the state would shard over GPUs is a design note, "
                                   "not a code path)"}))
        sys.exit(2)
    if args.workload == "adjoint":
        out = run_adjoint(args)
    elif args.workload == "hea":
        out = run_hea(args)
    elif args.workload == "ucc":
        out = run_ucc(args)
    elif args.workload == "serve":
        out = run_serve(args)
    elif args.workload == "sample":
        out = run_sample(args)
    elif args.workload == "mfma":
        out = run_mfma(args)
    else:
        out = run_brickwork(args)
        if not args.no_extras:
            # the rest of the metric: config 3 adjoint vs plain, config 5 serving
            adj = run_adjoint(args, steps=20, warmup=3, n=22, layers=12)
            out["adjoint_22q"] = {k: adj[k] for k in (
                "value", "unit", "ms_per_step", "plain_path", "speedup_vs_plain",
                "speedup_vs_plain_gpu_only", "vs_complex128", "roofline", "config")}
            out["serve"] = serve_at(args.serve_rate, args.requests, args.slots, args.seed)
        out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_brickwork(
            out["config"]["qubits"], out["config"]["layers"], args.terms)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
