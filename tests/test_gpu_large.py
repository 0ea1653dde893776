"""States above 32 qubits (64-bit indices in every kernel): a 33-qubit state
(64 GiB of complex64) against closed forms.  RY(theta_q) on every qubit
followed by a CZ chain keeps the Z statistics of the product state and turns
X_q into X_q Z_{q-1} Z_{q+1}:
  <Z_q> = cos t_q,  <Z_q Z_q+1> = cos t_q cos t_q+1,
  <X_q> = sin t_q * prod over chain neighbours r of cos t_r,  <Y_q> = 0.
Qubit 0 is index bit 32."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 33


@pytest.fixture(scope="module")
def big_state(gpu):
    import torch

    from tyxonq_amd import _native as nat

    free, _ = torch.cuda.mem_get_info()
    if free < (8 << N) + (8 << 30):
        pytest.skip(f"needs {8 << N >> 30} GiB of free device memory")
    psi = torch.empty(1 << N, dtype=torch.complex64, device=nat.device())
    yield psi
    del psi
    torch.cuda.empty_cache()


@pytest.mark.parametrize("jit", [1, -1])
def test_33_qubit_closed_forms(big_state, jit):
    import tyxonq_amd as tq
    from tyxonq_amd import _native as nat
    from tyxonq_amd.engine import lower
    from tyxonq_amd.pauli import PauliSum

    rng = np.random.default_rng(33)
    th = rng.uniform(0.2, 2 * np.pi - 0.2, N)
    c = tq.Circuit(N)
    for q in range(N):
        c.ry(q, float(th[q]))
    for q in range(N - 1):
        c.cz(q, q + 1)
    L, _, _ = lower(c)
    plan = nat.Plan(N, L.segments[0][1], jit=jit)
    print(f"33q jit={jit}: {plan.info()}", flush=True)
    nat.init_basis(big_state, 0)
    plan.run(big_state)

    terms, want = [], []
    for q in range(N):
        terms.append((1.0, [("Z", q)]))
        want.append(np.cos(th[q]))
    for q in range(N - 1):
        terms.append((1.0, [("Z", q), ("Z", q + 1)]))
        want.append(np.cos(th[q]) * np.cos(th[q + 1]))
    for q in (0, 1, 16, 31, 32):
        nb = [r for r in (q - 1, q + 1) if 0 <= r < N]
        terms.append((1.0, [("X", q)]))
        want.append(np.sin(th[q]) * np.prod([np.cos(th[r]) for r in nb]))
        terms.append((1.0, [("Y", q)]))
        want.append(0.0)
    x, z, cf = PauliSum(N, terms).to_masks()
    ham = nat.Ham(N, x, z, cf)
    e, norm, per = ham.expval(big_state, per_term=True)
    assert abs(norm - 1.0) < 2e-5
    np.testing.assert_allclose(per, want, atol=2e-5)
    assert abs(e - float(np.sum(want))) < 1e-3


def test_33_qubit_sampling_marginals(big_state):
    import tyxonq_amd as tq
    from tyxonq_amd import _native as nat
    from tyxonq_amd.engine import lower

    rng = np.random.default_rng(33)
    th = rng.uniform(0.2, 2 * np.pi - 0.2, N)
    c = tq.Circuit(N)
    for q in range(N):
        c.ry(q, float(th[q]))
    for q in range(N - 1):
        c.cz(q, q + 1)
    L, _, _ = lower(c)
    nat.init_basis(big_state, 0)
    nat.Plan(N, L.segments[0][1]).run(big_state)
    shots = 200000
    idx, cnt = nat.sample_counts(big_state, shots, 7)
    assert int(cnt.sum()) == shots and np.all(np.diff(idx.astype(np.int64)) > 0)
    assert int(idx.max()) < (1 << N)
    for q in (0, 1, 15, 31, 32):
        bit = np.uint64(N - 1 - q)
        ones = float(cnt[((idx >> bit) & np.uint64(1)).astype(bool)].sum()) / shots
        p = np.sin(th[q] / 2) ** 2
        assert abs(ones - p) < 5 * np.sqrt(p * (1 - p) / shots) + 1e-9, (q, ones, p)
