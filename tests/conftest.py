"""Shared test configuration.

Markers:
  gpu - needs a HIP device and the built libtqsv.so; run with `-m gpu`.
CPU tests (`-m "not gpu"`) cover the oracle against the reference's known
answers, the host-side logic, and that the C ABI library loads and exports
every symbol of include/tqsv.h.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU and the built libtqsv.so")


def load_known_answers():
    return json.loads((GOLDEN / "reference_known_answers.json").read_text())["known_answers"]


def cplx(v):
    """[re, im] (possibly nested) -> complex numpy array."""
    import numpy as np

    a = np.asarray(v, dtype=float)
    return a[..., 0] + 1j * a[..., 1]


@pytest.fixture(scope="session")
def gpu():
    """The GPU engine; fails (does not skip) when the native library is absent."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a host without a visible HIP device")
    from tyxonq_amd import _native

    _native.load()
    return torch.device("cuda", 0)
