"""Writes tests/golden/fullsize_c128.json: complex128 results of the oracle at
the full benchmark sizes, which the reference itself cannot reach (its einsum
spec runs out of letters for n >= 23, libs/quantum_library/kernels/
statevector.py:50-55, and a 30-qubit complex128 state does not fit twice in
host memory).  Run once in the build container (CPU only, hours at 30 qubits):

    python tests/golden/make_fullsize_goldens.py [--only brickwork30,adjoint22,adjoint26,ucc20]

Computed with the complex128 C restatement (oracle/sv128.c through
oracle/c128.py, pinned by tests/test_oracle_c128.py), gate by gate in the
reference's order:

  brickwork30  30 qubits x 20 layers of RX.RZ + CZ brickwork (seed 1234) and
               the 500-term random 2-local Pauli sum: <P_k> per term, E, norm.
  adjoint22    22-qubit RY.RZ + CX-ring ansatz, 12 layers, 528 parameters
               ~ N(0, 0.5^2) (seed 1234), periodic Heisenberg chain + 0.5 sum Z:
               E and the full gradient by the reverse sweep, plus a sample of
               24 components by the parameter-shift rule (common.py:11-23).
  adjoint26    the same ansatz at 26 qubits, 2 layers, 104 parameters ~ N(0, 0.5^2)
               (default_rng(7)): E and the full gradient by the reverse sweep.
  ucc20        20 spin-orbitals, 10 electrons, synthetic integrals (seed 1234),
               UCCSD with spin-shared parameters, theta ~ N(0, 0.05^2) from
               default_rng(1234): E and the full gradient by the reverse sweep,
               plus 6 components by central differences (h = 1e-4).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import c128 as C  # noqa: E402
from tyxonq_amd import bench_workloads as W  # noqa: E402
from tyxonq_amd import chem  # noqa: E402

OUT = Path(__file__).resolve().parent / "fullsize_c128.json"
SEED = 1234


def log(msg: str) -> None:
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def brickwork30() -> dict:
    n, layers, nterms = 30, 20, 500
    c = W.brickwork_circuit(n, layers, seed=SEED)
    terms = W.random_2local_terms(n, nterms, seed=SEED)
    gates = C.gates_of(c.ops, n)
    t0 = time.time()
    psi = C.basis(n)
    for k, g in enumerate(gates):
        C.apply_gate(psi, g, n)
        if k % 100 == 0:
            log(f"brickwork30 gate {k}/{len(gates)}")
    nrm = C.norm2(psi, n)
    vals = []
    for k, (_, fac) in enumerate(terms):
        x, z = C.masks(fac, n)
        vals.append(C.pauli_dot(psi, psi, x, z, n).real)
        if k % 100 == 0:
            log(f"brickwork30 term {k}/{nterms}")
    coeff = np.array([cf for cf, _ in terms])
    e = float(np.dot(coeff, vals))
    return {"qubits": n, "layers": layers, "seed": SEED, "pauli_terms": nterms,
            "energy": e, "norm2": nrm, "term_values": [float(v) for v in vals],
            "cpu_s": round(time.time() - t0, 1)}


def adjoint22() -> dict:
    n, layers = 22, 12
    c, th = W.ryrz_ring_ansatz(n, layers, seed=SEED)
    th = np.asarray(th, dtype=float)
    terms = W.heisenberg_ring_terms(n, 0.5)
    t0 = time.time()
    gates = C.gates_of(c.ops, n, param_of=[(k, 1.0) for k in range(th.size)])
    e, grad, nrm = C.energy_and_grad(n, gates, terms, th.size)
    log(f"adjoint22 E={e:.12f} |g|={np.linalg.norm(grad):.6f}")
    sample = [int(i) for i in np.random.default_rng(SEED + 7).choice(th.size, 24, replace=False)]

    def energy(p):
        cc, _ = W.ryrz_ring_ansatz(n, layers, params=list(p))
        psi = C.run_gates(C.basis(n), C.gates_of(cc.ops, n), n)
        return float(np.dot([cf for cf, _ in terms], C.term_values(psi, terms, n)))

    ps = []
    for i in sample:
        p, m = th.copy(), th.copy()
        p[i] += 0.5 * np.pi
        m[i] -= 0.5 * np.pi
        ps.append(0.5 * (energy(p) - energy(m)))
        log(f"adjoint22 parameter shift {i}: {ps[-1]:.12f} vs sweep {grad[i]:.12f}")
    return {"qubits": n, "layers": layers, "seed": SEED, "params": int(th.size),
            "hamiltonian": "periodic Heisenberg XX+YY+ZZ + 0.5 sum Z", "energy": float(e),
            "norm2": nrm, "grad": [float(v) for v in grad], "ps_indices": sample,
            "ps_grad": [float(v) for v in ps], "cpu_s": round(time.time() - t0, 1)}


def adjoint26() -> dict:
    # two layers at 26 qubits: 8192 tiles, so a dual pass runs several tiles
    # per work-group; parameters from default_rng(7), as the test draws them
    n, layers = 26, 2
    th = np.random.default_rng(7).normal(0.0, 0.5, size=2 * n * layers)
    c, _ = W.ryrz_ring_ansatz(n, layers, params=list(th))
    terms = W.heisenberg_ring_terms(n, 0.5)
    t0 = time.time()
    gates = C.gates_of(c.ops, n, param_of=[(k, 1.0) for k in range(th.size)])
    e, grad, nrm = C.energy_and_grad(n, gates, terms, th.size)
    log(f"adjoint26 E={e:.12f} |g|={np.linalg.norm(grad):.6f}")
    return {"qubits": n, "layers": layers, "seed": 7, "params": int(th.size),
            "hamiltonian": "periodic Heisenberg XX+YY+ZZ + 0.5 sum Z", "energy": float(e),
            "norm2": nrm, "grad": [float(v) for v in grad], "cpu_s": round(time.time() - t0, 1)}


def ucc_setup(ns: int = 10, seed: int = SEED):
    no = ns // 2
    n = 2 * ns
    h, g = chem.synthetic_integrals(ns, seed)
    terms = chem.molecular_hamiltonian(h, g)
    ex, ids = chem.uccsd_excitations(no, ns - no)
    npar = max(ids) + 1
    th = np.random.default_rng(seed).normal(0.0, 0.05, size=npar)
    return n, no, terms, ex, ids, th


def ucc20() -> dict:
    n, no, terms, ex, ids, th = ucc_setup()
    t0 = time.time()
    param_of = [(pid, -2.0 * v.imag) for f_idx, pid in zip(ex, ids)
                for v, _, _ in chem.excitation_strings(f_idx, n)]

    def gates_for(p):
        ops = [tuple(op) for op in chem.ucc_circuit(list(p), n, (no, no), ex, ids).ops]
        return C.gates_of(ops, n, param_of=param_of)

    e, grad, nrm = C.energy_and_grad(n, gates_for(th), terms, th.size)
    log(f"ucc20 E={e:.12f} |g|={np.linalg.norm(grad):.6f} max|g|={np.abs(grad).max():.4f}")
    order = np.argsort(-np.abs(grad))
    sample = [int(i) for i in order[:3]] + [int(i) for i in
                                            np.random.default_rng(SEED + 9).choice(th.size, 3,
                                                                                   replace=False)]
    coeff = np.array([cf for cf, _ in terms])

    def energy(p):
        psi = C.run_gates(C.basis(n), gates_for(p), n)
        return float(np.dot(coeff, C.term_values(psi, terms, n)))

    fd = []
    h = 1e-4
    for i in sample:
        p, m = th.copy(), th.copy()
        p[i] += h
        m[i] -= h
        fd.append((energy(p) - energy(m)) / (2 * h))
        log(f"ucc20 central difference {i}: {fd[-1]:.10f} vs sweep {grad[i]:.10f}")
    return {"spin_orbitals": n, "electrons": 2 * no, "seed": SEED, "params": int(th.size),
            "excitations": len(ex), "pauli_terms": len(terms), "theta_sigma": 0.05,
            "energy": float(e), "norm2": nrm, "grad": [float(v) for v in grad],
            "fd_indices": sample, "fd_h": h, "fd_grad": [float(v) for v in fd],
            "cpu_s": round(time.time() - t0, 1)}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="adjoint22,ucc20,brickwork30")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 8)
    args = ap.parse_args()
    C.set_threads(args.threads)
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    data["_about"] = ("complex128 oracle results at full benchmark sizes; written by "
                      "tests/golden/make_fullsize_goldens.py")
    for name in args.only.split(","):
        log(f"start {name}")
        data[name] = {"adjoint22": adjoint22, "adjoint26": adjoint26, "ucc20": ucc20,
                      "brickwork30": brickwork30}[name]()
        OUT.write_text(json.dumps(data, indent=1) + "\n")
        log(f"wrote {name} to {OUT}")


if __name__ == "__main__":
    main()
