"""GPU shot sampling (tqsv_sampler_*) against the oracle's restatement of the
same inverse-CDF sampler: identical counts for the same state, shot count and
seed; size-independent properties at 26 qubits; error paths."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import sampling as S
from tests.gpu_util import to_gpu

pytestmark = pytest.mark.gpu


def _state(rng, n, zero_frac=0.2):
    v = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    v[rng.random(1 << n) < zero_frac] = 0
    return v


CASES = [(1, 10, 0), (3, 1, 5), (5, 7, 2), (8, 1000, 7), (12, 5000, 1), (13, 200000, 2),
         (14, 3, 4), (17, 100000, 3), (20, 300000, 9), (22, 20000, 6), (24, 1000, 3)]


@pytest.mark.parametrize("n,shots,seed", CASES)
def test_counts_equal_oracle(gpu, n, shots, seed):
    from tyxonq_amd import _native as nat

    rng = np.random.default_rng(100 + n)
    v = _state(rng, n)
    s = nat.Sampler(n, max(shots, 16))
    idx, cnt = s.sample(to_gpu(v), shots, seed)
    widx, wcnt = S.sample_counts(v, shots, seed)
    np.testing.assert_array_equal(idx, widx)
    np.testing.assert_array_equal(cnt, wcnt)
    assert int(cnt.sum()) == shots
    s.close()


def test_sparse_states_and_zero_mass_blocks(gpu):
    from tyxonq_amd import _native as nat

    n = 20
    ghz = np.zeros(1 << n, dtype=np.complex64)
    ghz[0] = ghz[-1] = np.sqrt(0.5)
    idx, cnt = nat.sample_counts(to_gpu(ghz), 50000, seed=8)
    assert idx.tolist() == [0, (1 << n) - 1] and int(cnt.sum()) == 50000
    widx, wcnt = S.sample_counts(ghz, 50000, seed=8)
    np.testing.assert_array_equal(cnt, wcnt)
    # mass in two amplitudes of one late block, every other block empty
    v = np.zeros(1 << n, dtype=np.complex64)
    v[(1 << n) - 4096 + 17] = 0.3
    v[(1 << n) - 4096 + 4000] = 0.4j
    idx, cnt = nat.sample_counts(to_gpu(v), 777, seed=1)
    widx, wcnt = S.sample_counts(v, 777, seed=1)
    np.testing.assert_array_equal(idx, widx)
    np.testing.assert_array_equal(cnt, wcnt)
    one = np.zeros(1 << n, dtype=np.complex64)
    one[12345] = 1
    idx, cnt = nat.sample_counts(to_gpu(one), 10, seed=2)
    assert idx.tolist() == [12345] and cnt.tolist() == [10]


def test_large_state_properties(gpu):
    # 26 qubits, 10^6 shots: counts sum to the shots, indices ascend, and the
    # marginal of every one of the top 4 index bits matches <Z> within 6 sigma
    import torch

    from tyxonq_amd import _native as nat

    n, shots = 26, 1_000_000
    g = torch.Generator(device="cuda").manual_seed(5)
    psi = torch.complex(torch.randn(1 << n, device="cuda", generator=g),
                        torch.randn(1 << n, device="cuda", generator=g)).to(torch.complex64)
    norm2 = nat.norm2(psi)
    idx, cnt = nat.sample_counts(psi, shots, seed=77)
    assert int(cnt.sum()) == shots
    assert np.all(np.diff(idx.astype(np.int64)) > 0)
    for bit in range(n - 4, n):
        p1 = (1.0 - nat.expect_z(psi, bit) / norm2) / 2.0
        got = cnt[((idx >> np.uint64(bit)) & np.uint64(1)).astype(bool)].sum() / shots
        assert abs(got - p1) < 6 * np.sqrt(p1 * (1 - p1) / shots)


def test_sampler_errors(gpu):
    from tyxonq_amd import _native as nat

    s = nat.Sampler(6, 100)
    with pytest.raises(ValueError):
        s.sample(to_gpu(np.zeros(64, dtype=np.complex64)), 10, 0)
    with pytest.raises(ValueError):
        s.sample(to_gpu(np.ones(64, dtype=np.complex64)), 101, 0)
    idx, cnt = s.sample(to_gpu(np.ones(64, dtype=np.complex64)), 0, 0)
    assert idx.size == 0 and cnt.size == 0
    with pytest.raises(ValueError):
        s.sample(to_gpu(np.ones(128, dtype=np.complex64)), 10, 0)
    s.close()


def test_engine_counts_from_gpu_sampler(gpu):
    import tyxonq_amd as tq
    from tyxonq_amd.engine import StatevectorEngine

    c = tq.Circuit(5).h(0).h(1).cx(1, 2).ry(3, 0.7).measure_z(0).measure_z(1)
    eng = StatevectorEngine()
    out = eng.run(c, shots=4000, seed=21)
    psi = eng.last_state.detach().cpu().numpy()
    widx, wcnt = S.sample_counts(psi, 4000, 21)
    assert out["result"] == {format(int(i), "05b"): int(k) for i, k in zip(widx, wcnt)}
