"""Build recipe for libtqsv.so (gfx950 HIP kernels + C++ host runtime).

The library is built in-tree (tyxonq_amd/libtqsv.so) so that it travels with
the source tree.  HIP sources go through hipcc for gfx950; the host-only C++
(planner, C ABI) goes through g++ against the HIP runtime headers.

    python -m tyxonq_amd._build          # build (always recompiles)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libtqsv.so"
OBJ = PKG / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # no packed FP32 instructions in device code
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-result",
]
CXX_FLAGS = [
    "-O2",
    "-std=c++17",
    "-fPIC",
    "-D__HIP_PLATFORM_AMD__",
    f"-I{ROCM / 'include'}",
    "-Wall",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(
            "command failed: " + " ".join(cmd) + "\n" + res.stdout[-4000:] + res.stderr[-8000:]
        )


def sources() -> tuple[list[Path], list[Path]]:
    hip = sorted(CSRC.glob("*.hip"))
    cpp = sorted(CSRC.glob("*.cpp"))
    return hip, cpp


def _stale(src: Path, obj: Path, headers: list[Path]) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return src.stat().st_mtime > t or any(h.stat().st_mtime > t for h in headers)


def build(verbose: bool = False, incremental: bool = False) -> Path:
    """Compile every source (or, with incremental=True, only those older than
    their object or any header) and link libtqsv.so."""
    hipcc = str(ROCM / "bin" / "hipcc")
    OBJ.mkdir(exist_ok=True)
    hip, cpp = sources()
    headers = sorted(CSRC.glob("*.h")) + sorted((PKG.parent / "include").glob("*.h"))
    jobs = []
    for src in hip:
        obj = OBJ / (src.stem + ".o")
        if not incremental or _stale(src, obj, headers):
            jobs.append([hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in cpp:
        obj = OBJ / (src.stem + ".o")
        if not incremental or _stale(src, obj, headers):
            jobs.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    if jobs:
        with ThreadPoolExecutor(max_workers=workers) as ex:
            for cmd, _ in zip(jobs, ex.map(_run, jobs)):
                if verbose:
                    print(" ".join(cmd[:1] + cmd[-3:]))
    objs = [str(OBJ / (s.stem + ".o")) for s in hip + cpp]
    tmp = OUT.with_name(OUT.name + ".tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv)
    print(p)
