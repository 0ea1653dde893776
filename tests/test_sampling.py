"""Shot sampling oracle (CPU): the Philox stream and the inverse-CDF sampler.

The GPU sampler is checked against this restatement count for count
(tests/test_gpu_sampling.py); here the restatement itself is pinned.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import stats

from oracle import sampling as S


@pytest.mark.parametrize("key", [(0, 0), (7, 9), (2 ** 64 - 1, 12345), (0x243F6A8885A308D3, 1)])
def test_philox_rounds_match_numpy_philox(key):
    # numpy's Philox is an independent implementation of Philox4x64-10; it
    # increments its 256-bit counter before each block
    for c0 in (0, 1, 5, 2 ** 63, 2 ** 64 - 2):
        bg = np.random.Philox(key=np.array(key, dtype=np.uint64),
                              counter=np.array([c0, 0, 0, 0], dtype=np.uint64))
        raw = [int(v) for v in bg.random_raw(8)]
        for k in (1, 2):  # the counter is a 256-bit integer (word 0 least significant)
            c = c0 + k
            words = [(c >> (64 * i)) & (2 ** 64 - 1) for i in range(4)]
            assert raw[4 * (k - 1):4 * k] == S.philox4x64_10(words, key)


def test_stream_starts_at_counter_zero():
    w = S.stream_words(42, 9)
    assert [int(v) for v in w[:4]] == S.philox4x64_10([0, 0, 0, 0], [42, 0])
    assert [int(v) for v in w[4:8]] == S.philox4x64_10([1, 0, 0, 0], [42, 0])
    assert int(w[8]) == S.philox4x64_10([2, 0, 0, 0], [42, 0])[0]
    u = S.uniforms(42, 9)
    assert np.all((u >= 0) & (u < 1))
    assert u[0] == (S.philox4x64_10([0, 0, 0, 0], [42, 0])[0] >> 11) * 2.0 ** -53


def test_single_amplitude_and_zero_mass_blocks():
    n = 14  # 4 sampling blocks; the mass sits in the last one
    psi = np.zeros(1 << n, dtype=np.complex64)
    psi[(1 << n) - 5] = 0.6 + 0.8j
    idx, cnt = S.sample_counts(psi, 1000, seed=3)
    assert idx.tolist() == [(1 << n) - 5] and cnt.tolist() == [1000]
    psi[:] = 0
    psi[7] = 1.0
    idx, cnt = S.sample_counts(psi, 10, seed=3)
    assert idx.tolist() == [7] and cnt.tolist() == [10]
    with pytest.raises(ValueError):
        S.sample_counts(np.zeros(8, dtype=np.complex64), 5, seed=1)


@pytest.mark.parametrize("n,shots", [(3, 20000), (9, 200000), (13, 300000)])
def test_distribution_chi_square(n, shots):
    rng = np.random.default_rng(n)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi[rng.random(1 << n) < 0.2] = 0  # zero-probability states are never drawn
    p = S.probabilities(psi)
    p /= p.sum()
    idx, cnt = S.sample_counts(psi, shots, seed=11)
    assert int(cnt.sum()) == shots and np.all(np.diff(idx.astype(np.int64)) > 0)
    assert np.all(p[idx.astype(np.int64)] > 0)
    full = np.zeros(1 << n)
    full[idx.astype(np.int64)] = cnt
    # pool cells so that every expected count is >= 5
    order = np.argsort(p)
    exp_, obs_ = [], []
    acc_e = acc_o = 0.0
    for i in order:
        if p[i] == 0:
            continue
        acc_e += p[i] * shots
        acc_o += full[i]
        if acc_e >= 5:
            exp_.append(acc_e)
            obs_.append(acc_o)
            acc_e = acc_o = 0.0
    if acc_e > 0:
        exp_[-1] += acc_e
        obs_[-1] += acc_o
    chi2 = float(np.sum((np.array(obs_) - np.array(exp_)) ** 2 / np.array(exp_)))
    assert stats.chi2.sf(chi2, len(exp_) - 1) > 1e-4


def test_two_amplitudes_binomial():
    psi = np.zeros(1 << 5, dtype=np.complex64)
    psi[3] = 0.5
    psi[20] = np.sqrt(0.75)
    idx, cnt = S.sample_counts(psi, 100000, seed=5)
    assert idx.tolist() == [3, 20]
    # |z| of the binomial count below 5 sigma
    assert abs(cnt[0] - 25000) < 5 * np.sqrt(100000 * 0.25 * 0.75)
