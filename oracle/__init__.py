"""CPU oracle for the statevector path - TEST INFRASTRUCTURE ONLY.

A NumPy complex128 restatement of the reference's algorithm (TyxonQ,
src/tyxonq/libs/quantum_library/kernels/statevector.py, gates.py, pauli.py,
devices/simulators/statevector/engine.py, libs/quantum_library/kernels/common.py).
Each function cites the reference file:line it follows.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker or the timed CPU baseline; the
product path (tyxonq_amd) never calls it and has no CPU fallback.

How the oracle is pinned (see tests/test_oracle.py):
  * the reference's own known-answer tests, transcribed into
    tests/golden/reference_known_answers.json;
  * dense.py, an independent brute-force construction (full 2^n x 2^n
    Kronecker-product unitaries, n <= 10) for gate application, Pauli
    expectations and Z expectations;
  * closed forms: rotation periodicity and composition, <Z> after RX(t) =
    cos t, the exact parameter-shift rule for Pauli rotations, central
    finite differences, Jordan-Wigner anticommutation relations.
"""
