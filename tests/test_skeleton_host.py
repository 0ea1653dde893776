"""A plan update keeps its generated kernels when the new program, with its
value words (angles, phases, scales) zeroed, equals the one last compiled.
That is sound only if the generated source of a pass depends on nothing but
those non-value words.  tests/skeleton_check.cpp plans the same structure
with several parameter sets (small and large changes, angles of exactly pi)
and checks that equal skeletons always give equal sources (CPU only)."""
from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest

import tyxonq_amd as tq  # noqa: F401
from tyxonq_amd import _native as nat

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    lib = nat.lib_path()
    out = tmp_path_factory.mktemp("skel") / "skeleton_check"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = ["g++", "-O1", "-std=c++17", f"-I{rocm}/include", "-D__HIP_PLATFORM_AMD__",
           f"-I{os.path.join(ROOT, 'tyxonq_amd', 'csrc')}", os.path.join(HERE, "skeleton_check.cpp"),
           "-o", str(out), f"-L{lib.parent}", "-ltqsv", f"-Wl,-rpath,{lib.parent}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def _run(checker, n, dual, max_cost, files):
    r = subprocess.run([str(checker), str(n), "1" if dual else "0", str(max_cost), *map(str, files)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
    return r.stdout


def test_ucc_updates(checker, tmp_path):
    from tyxonq_amd import chem

    ns, no = 6, 3
    ex, ids = chem.uccsd_excitations(no, ns - no)
    ans = chem.UCCAnsatz(2 * ns, (no, no), ex, ids)
    rng = np.random.default_rng(5)
    base = rng.normal(0, 0.05, ans.nparams)
    sets = [base, base + rng.normal(0, 0.01, ans.nparams), rng.normal(0, 0.05, ans.nparams),
            rng.normal(0, 2.0, ans.nparams), np.full(ans.nparams, np.pi)]
    files = []
    for k, th in enumerate(sets):
        f = tmp_path / f"ucc{k}.ops"
        ans.ops(th).tofile(f)
        files.append(f)
    for dual in (False, True):
        _run(checker, 2 * ns, dual, 100000, files)


def test_layered_ansatz_updates(checker, tmp_path):
    import torch

    from tyxonq_amd import bench_workloads as W
    from tyxonq_amd.engine import lower

    n, layers = 14, 4
    files = []
    for k, sc in enumerate([0.5, 0.5, 0.5, 3.0, None]):
        th = np.random.default_rng(100 + k).normal(0, sc or 1.0, 2 * n * layers)
        if sc is None:
            th[:] = np.pi
        tt = torch.tensor(th, dtype=torch.float64, requires_grad=True)
        circ, _ = W.ryrz_ring_ansatz(n, layers, params=[tt[i] for i in range(tt.numel())])
        L, _, _ = lower(circ, track_grad=True)
        f = tmp_path / f"adj{k}.ops"
        np.asarray(L.segments[0][1]).tofile(f)
        files.append(f)
    for dual in (False, True):
        out = _run(checker, n, dual, 0, files)
        assert "file 1: skeleton equal" in out  # a small change keeps the kernels
