"""Writes reference_known_answers.json: the known-answer values that the
reference's own tests assert on this path, transcribed as data.

Every entry names the reference test (file:line) it comes from.  The values
are literal numbers from those tests (no reference code is imported or run).
Circuits are op-tuple lists in the reference's IR format.

    python tests/golden/make_known_answers.py
"""
from __future__ import annotations

import json
import math
from pathlib import Path

OUT = Path(__file__).with_name("reference_known_answers.json")


def c(z: complex) -> list:
    return [float(complex(z).real), float(complex(z).imag)]


def main() -> None:
    s2 = 1 / math.sqrt(2)
    ka = []
    # --- final states -------------------------------------------------------
    ka.append({
        "name": "sqrt_x_unitary",
        "source": "tests_core_module/test_circuit_unitary.py:16-26",
        "n": 1,
        "ops": [["unitary", 0, "_u0"]],
        "unitaries": {"_u0": [[c(0.5 + 0.5j), c(0.5 - 0.5j)], [c(0.5 - 0.5j), c(0.5 + 0.5j)]]},
        "state": [c(0.5 + 0.5j), c(0.5 - 0.5j)],
        "atol": 1e-8,
    })
    ka.append({
        "name": "iswap_on_10",
        "source": "tests_core_module/test_circuit_unitary.py:29-44",
        "n": 2,
        "ops": [["x", 0], ["unitary", 0, 1, "_u1"]],
        "unitaries": {"_u1": [[c(1), c(0), c(0), c(0)], [c(0), c(0), c(1j), c(0)],
                              [c(0), c(1j), c(0), c(0)], [c(0), c(0), c(0), c(1)]]},
        "state": [c(0), c(1j), c(0), c(0)],
        "atol": 1e-8,
    })
    ka.append({
        "name": "ry_pi_over_4",
        "source": "tests_core_module/test_circuit_unitary.py:47-62",
        "n": 1,
        "ops": [["ry", 0, math.pi / 4]],
        "state": [c(math.cos(math.pi / 8)), c(math.sin(math.pi / 8))],
        "atol": 1e-8,
    })
    ka.append({
        "name": "hh_then_iswap",
        "source": "tests_core_module/test_circuit_unitary.py:80-102",
        "n": 2,
        "ops": [["h", 0], ["h", 1], ["iswap", 0, 1]],
        "state": [c(0.5), c(0.5j), c(0.5j), c(0.5)],
        "atol": 1e-8,
    })
    ka.append({
        "name": "h_on_q0_of_3",
        "source": "tests_core_module/test_circuit_unitary.py:140-160",
        "n": 3,
        "ops": [["h", 0]],
        "state": [c(s2), c(0), c(0), c(0), c(s2), c(0), c(0), c(0)],
        "atol": 1e-8,
    })
    # --- expectations -------------------------------------------------------
    ka.append({
        "name": "bell_z_expectations",
        "source": "tests_core_module/test_devices_simulators_gates.py:17-30",
        "n": 2,
        "ops": [["h", 0], ["cx", 0, 1]],
        "z": {"0": 0.0, "1": 0.0},
        "norm": 1.0,
        "atol": 1e-12,
    })
    ka.append({
        "name": "plus_state_probs",
        "source": "tests_core_module/test_statevector_engine_probs.py:7-27",
        "n": 2,
        "ops": [["h", 0]],
        "prob_sum_indices": [[0, 2], 1.0],
        "amplitudes": {"01": 0.0},
        "atol": 1e-8,
    })
    ka.append({
        "name": "x_on_plus",
        "source": "tests_core_module/test_circuit_expectation.py:9-17",
        "n": 1,
        "ops": [["h", 0]],
        "pauli_terms": [[1.0, [["X", 0]]]],
        "energy": 1.0,
        "atol": 1e-10,
    })
    ka.append({
        "name": "z_on_zero",
        "source": "tests_core_module/test_circuit_expectation.py:20-27",
        "n": 1,
        "ops": [],
        "pauli_terms": [[1.0, [["Z", 0]]]],
        "energy": 1.0,
        "atol": 1e-10,
    })
    ka.append({
        "name": "bell_zz",
        "source": "tests_core_module/test_circuit_expectation.py:30-39",
        "n": 2,
        "ops": [["h", 0], ["cx", 0, 1]],
        "pauli_terms": [[1.0, [["Z", 0], ["Z", 1]]]],
        "energy": 1.0,
        "atol": 1e-10,
    })
    ka.append({
        "name": "bell_xx",
        "source": "tests_core_module/test_circuit_expectation.py:42-51",
        "n": 2,
        "ops": [["h", 0], ["cx", 0, 1]],
        "pauli_terms": [[1.0, [["X", 0], ["X", 1]]]],
        "energy": 1.0,
        "atol": 1e-10,
    })
    ka.append({
        "name": "tfi_plus_plus_plus",
        "source": "tests_core_module/test_circuit_expectation.py:54-69",
        "n": 3,
        "ops": [["h", 0], ["h", 1], ["h", 2]],
        "pauli_terms": [[-1.0, [["X", 0]]], [-1.0, [["X", 1]]], [-1.0, [["X", 2]]],
                        [1.0, [["Z", 0], ["Z", 1]]], [1.0, [["Z", 1], ["Z", 2]]]],
        "energy": -3.0,
        "atol": 1e-10,
    })
    ka.append({
        "name": "rx_half_z",
        "source": "tests_core_module/test_circuit_state_autograd.py:78-102",
        "n": 1,
        "ops": [["rx", 0, 0.5]],
        "pauli_terms": [[1.0, [["Z", 0]]]],
        "energy": math.cos(0.5),
        "atol": 1e-6,
    })
    # --- gradients ----------------------------------------------------------
    ka.append({
        "name": "rx_zero_gradient",
        "source": "tests_core_module/test_circuit_state_autograd.py:154-189",
        "n": 1,
        "ops": [["rx", 0, "p0"]],
        "params": [0.0],
        "pauli_terms": [[1.0, [["Z", 0]]]],
        "grad": [0.0],
        "atol": 1e-6,
    })
    # --- run() expectations -------------------------------------------------
    ka.append({
        "name": "project_and_reset",
        "source": "tests_core_module/test_mid_measure_postselect.py:9-16",
        "n": 2,
        "ops": [["h", 0], ["project_z", 0, 0], ["reset", 1], ["measure_z", 0], ["measure_z", 1]],
        "run_expectations": {"Z0": 1.0, "Z1": 1.0},
        "atol": 1e-8,
    })
    ka.append({
        "name": "rx_pi_z",
        "source": "tests_core_module/test_noise_integration.py:9-20",
        "n": 1,
        "ops": [["rx", 0, math.pi], ["measure_z", 0]],
        "run_expectations": {"Z0": -1.0},
        "atol": 1e-8,
    })
    ka.append({
        "name": "rx_pi_z_depolarizing",
        "source": "tests_core_module/test_noise_integration.py:9-20",
        "n": 1,
        "ops": [["rx", 0, math.pi], ["measure_z", 0]],
        "noise": {"type": "depolarizing", "p": 0.3},
        "run_expectations": {"Z0": -(1.0 - 4.0 * 0.3 / 3.0)},
        "atol": 1e-6,
    })
    # --- matrices -----------------------------------------------------------
    ka.append({
        "name": "pauli_x_matrix",
        "source": "tests_core_module/test_core_operations_pauli.py:22-25",
        "codes": [[1]],
        "matrix": [[c(0), c(1)], [c(1), c(0)]],
    })
    ka.append({
        "name": "zz_matrix",
        "source": "tests_core_module/test_core_operations_pauli.py:28-31",
        "codes": [[3, 3]],
        "matrix": [[c(1), c(0), c(0), c(0)], [c(0), c(-1), c(0), c(0)],
                   [c(0), c(0), c(-1), c(0)], [c(0), c(0), c(0), c(1)]],
    })
    ka.append({
        "name": "controlled_x_is_cx",
        "source": "tests_core_module/test_devices_simulators_gates_extra.py:42-57",
        "controlled": {"U": [[c(0), c(1)], [c(1), c(0)]], "num_controls": 1},
        "matrix": [[c(1), c(0), c(0), c(0)], [c(0), c(1), c(0), c(0)],
                   [c(0), c(0), c(0), c(1)], [c(0), c(0), c(1), c(0)]],
    })
    OUT.write_text(json.dumps({"known_answers": ka}, indent=1))
    print(f"wrote {OUT} ({len(ka)} entries)")


if __name__ == "__main__":
    main()
