"""The C ABI library loads and exports what include/tqsv.h declares (CPU only).

No compute call is made here: there is no GPU.  tqsv_plan_inspect is host
only, so the planner is exercised on the CPU through it.
"""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tyxonq_amd import _native as nat

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "tqsv.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(tqsv_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_header_symbols():
    lib = nat.load()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), f"{name} declared in tqsv.h but not exported"
    assert set(names) == set(nat.exported_symbols()), "ctypes signature table out of sync with header"
    assert nat.version() == 100
    assert lib.tqsv_last_error() is not None


def test_op_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "tqsv.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(tqsv_op),"
        " offsetof(tqsv_op, cmask), offsetof(tqsv_op, theta), offsetof(tqsv_op, gcoef),"
        " offsetof(tqsv_op, m), sizeof(tqsv_plan_info)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    size, off_cmask, off_theta, off_gcoef, off_m, info_size = map(int, out)
    d = nat.OP_DTYPE
    assert size == d.itemsize
    assert off_cmask == d.fields["cmask"][1]
    assert off_theta == d.fields["theta"][1]
    assert off_gcoef == d.fields["gcoef"][1]
    assert off_m == d.fields["m"][1]
    assert info_size == ctypes.sizeof(nat.PlanInfo)


def _ops_brickwork(n, layers, seed=1234):
    from tyxonq_amd import bench_workloads as W

    c = W.brickwork_circuit(n, layers, seed)
    from tyxonq_amd.engine import lower

    L, _, _ = lower(c)
    return L.segments[0][1]


def test_planner_inspect_brickwork_passes():
    ops = _ops_brickwork(20, 4)
    info = nat.inspect(20, ops)
    assert info["n"] == 20
    assert info["ops_in"] == len(ops)
    # RX.RZ pairs merge into one 2x2 per qubit and layer; CZs stay
    assert info["gates_after_merge"] < info["ops_in"]
    assert 1 <= info["passes"] <= 4 * 20
    assert info["bytes_per_run"] == info["passes"] * 16 * (1 << 20)
    unfused = nat.inspect(20, ops, fuse=False)
    assert unfused["passes"] == unfused["gates_after_merge"]
    assert info["passes"] < unfused["passes"]


def test_planner_rejects_bad_ops():
    ops = np.zeros(1, dtype=nat.OP_DTYPE)
    ops[0]["kind"] = nat.OP_MAT1
    ops[0]["t"] = (7, -1)
    with pytest.raises(ValueError, match="out of range"):
        nat.inspect(4, ops)
    ops[0]["t"] = (1, -1)
    ops[0]["cmask"] = 2
    with pytest.raises(ValueError, match="control"):
        nat.inspect(4, ops)
    ops[0]["cmask"] = 0
    ops[0]["kind"] = 9
    with pytest.raises(ValueError, match="unknown op kind"):
        nat.inspect(4, ops)
