#!/usr/bin/env python3
"""Benchmark of the MI355X statevector engine (one GPU).

Default workload: 30-qubit complex64 brickwork of 20 layers (RX.RZ on every
qubit, CZ on even/odd pairs) from |0...0>, then the norm and the energy of a
random 500-term 2-local Pauli sum, both reduced in FP64.  One step = that
whole job with the state resident in HBM.

Output: one JSON line.  `value` is the step's algorithmic HBM traffic over the
step time (GB/s): one state write for the initial state (8 B/amp), every gate
pass reads and writes the state (16 B/amp), every expectation pass reads it
(8 B/amp); passes are counted from the launch list.  `roofline` describes the
dominant kernel (the fused gate pass) with its own launch-averaged duration
measured by HIP events on the launch stream; `cpu_baseline` times the CPU
oracle (the reference's per-gate contraction restated in NumPy complex128)
on a bounded sample of the same 30-qubit workload.

    python bench.py [--steps K] [--warmup W] [--workload brickwork|adjoint|hea]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="brickwork", choices=["brickwork", "adjoint", "hea"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--layers", type=int, default=None)
    ap.add_argument("--terms", type=int, default=500)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--tile-bits", type=int, default=0)
    ap.add_argument("--max-cost", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one device copy of the state (HBM counter calibration)")
    ap.add_argument("--check", action="store_true",
                    help="compare the energy with a fused-vs-unfused run")
    return ap.parse_args()


def _sync():
    import torch

    torch.cuda.synchronize()


def _events():
    import torch

    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def cpu_baseline_brickwork(n: int, seed: int) -> dict:
    """Time the oracle's per-gate contraction (complex128) on three gates of
    the 30-qubit brickwork: RX on qubit 0, RZ on qubit n//2, CZ(n//2, n//2+1)."""
    from oracle import statevector as O

    try:
        from threadpoolctl import threadpool_info

        threads = max([d.get("num_threads", 1) for d in threadpool_info()] or [1])
    except Exception:
        threads = 1
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    samples = []
    t0 = time.perf_counter()
    psi = O.apply_1q(psi, O.gate_rx(0.3), 0, n)
    samples.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    psi = O.apply_1q(psi, O.gate_rz(0.7), n // 2, n)
    samples.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    psi = O.apply_2q(psi, O.gate_cz(), n // 2, n // 2 + 1, n)
    samples.append(time.perf_counter() - t0)
    del psi
    per_gate = float(np.mean(samples))
    gbps = 2 * 16 * (1 << n) / per_gate / 1e9
    return {
        "value": round(gbps, 3),
        "unit": "GB/s",
        "cores": int(threads),
        "kind": "port",
        "sample": (f"NumPy complex128 per-gate contraction (oracle restatement of the reference's "
                   f"apply_1q/apply_2q) on 3 gates of the {n}-qubit brickwork (RX q0, RZ q{n//2}, "
                   f"CZ q{n//2},{n//2+1}); value = 2*16*2^n B per gate / mean gate time"),
        "per_gate_s": round(per_gate, 3),
        "job_s_extrapolated_gates_only": round(per_gate * (20 * (2 * n) + 290), 1),
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
    plan = nat.Plan(n, ops, tile_bits=args.tile_bits, max_cost=args.max_cost)
    info = plan.info()
    x, z, c = PauliSum(n, terms).to_masks()
    ham = nat.Ham(n, x, z, c)
    dev = nat.device()
    psi = torch.empty(1 << n, dtype=torch.complex64, device=dev)
    n_exp_passes = None

    def step():
        nat.init_basis(psi, 0)
        plan.run(psi)
        return ham.expval(psi)

    for _ in range(args.warmup):
        step()
    _sync()
    # expectation pass count from the plan of the Hamiltonian: measured by timing below
    gate_ms, exp_ms = [], []
    energies = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        nat.init_basis(psi, 0)
        e0, e1 = _events()
        e0.record()
        plan.run(psi)
        e1.record()
        t0 = time.perf_counter()
        energies.append(ham.expval(psi))  # synchronizes the stream
        exp_ms.append((time.perf_counter() - t0) * 1e3)
        gate_ms.append(e0.elapsed_time(e1))
    _sync()
    t_total = time.perf_counter() - t_start
    ms_per_step = t_total / args.steps * 1e3
    dim = 1 << n
    gate_pass_bytes = 16 * dim
    passes = info["passes"]
    n_exp_passes = ham_passes(ham)
    step_bytes = 8 * dim + passes * gate_pass_bytes + n_exp_passes * 8 * dim
    value = step_bytes / (ms_per_step / 1e3) / 1e9
    avg_gate_ms = float(np.mean(gate_ms))
    per_launch_ms = avg_gate_ms / passes
    achieved = gate_pass_bytes / (per_launch_ms / 1e3) / 1e9
    energy, norm = energies[-1]
    ngates = len(ops)
    out = {
        "metric": "30q gate GB/s (% HBM peak); VQE energy+grad/s and p99 latency vs plain",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 re/im), fp64 reductions",
        "data": "synthetic (seeded brickwork circuit and random 2-local Pauli sum)",
        "config": {"workload": f"brickwork {n}q x {layers} layers + {args.terms}-term 2-local <H>",
                   "qubits": n, "layers": layers, "gates": ngates, "pauli_terms": args.terms,
                   "seed": args.seed, "gate_passes": passes, "expval_passes": n_exp_passes,
                   "tile_bits": info["tile_bits"], "phases": info["phases"],
                   "gates_after_merge": info["gates_after_merge"], "parallelism": "1 GPU"},
        "gate_ms": round(avg_gate_ms, 3),
        "gate_GBps": round(passes * gate_pass_bytes / (avg_gate_ms / 1e3) / 1e9, 1),
        "gate_equiv_GBps": round(info["gates_after_merge"] * gate_pass_bytes / (avg_gate_ms / 1e3) / 1e9, 1),
        "layer_ms": round(avg_gate_ms / layers, 3),
        "expval_ms": round(float(np.mean(exp_ms)), 3),
        "energy": energy,
        "norm2": norm,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": None,
            "kernel": "tqsv k_pass (fused tile pass)",
            "bytes_per_launch": gate_pass_bytes,
            "avg_launch_ms": round(per_launch_ms, 4),
        },
    }
    if args.check:
        ref = nat.Plan(n, ops, fuse=True, max_cost=100000)
        nat.init_basis(psi, 0)
        ref.run(psi)
        e2, _ = ham.expval(psi)
        out["check_energy_other_plan"] = e2
    if args.calib_copy:
        other = torch.empty_like(psi)
        other.copy_(psi)
        _sync()
        del other
    del psi
    torch.cuda.empty_cache()
    return out


def ham_passes(ham) -> int:
    """Read-only state sweeps of one expectation of a planned Hamiltonian."""
    return int(ham.info()["passes"])


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not a code path)"}))
        sys.exit(2)
    if args.workload != "brickwork":
        from tyxonq_amd import bench_modes

        out = bench_modes.run(args)
    else:
        out = run_brickwork(args)
    if not args.no_cpu_baseline and args.workload == "brickwork":
        out["cpu_baseline"] = cpu_baseline_brickwork(out["config"]["qubits"], args.seed)
    elif "cpu_baseline" not in out:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
