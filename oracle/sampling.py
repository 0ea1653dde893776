"""Shot sampling: NumPy restatement - TEST INFRASTRUCTURE ONLY (see
oracle/__init__.py).

The reference samples counts on the host: |psi|^2, optional noise mixing,
``choice`` with p = |psi|^2 / sum, ``bincount``, big-endian bitstrings
(src/tyxonq/devices/simulators/statevector/engine.py:377-465).  Its generator
is unseeded (``nb.rng(None)``, engine.py:380), so its counts are not
reproducible and there are no count-level golden vectors.  The GPU sampler
draws from the same distribution by inverse CDF with a counter-based stream,
and this module restates that algorithm so that the GPU counts can be
checked exactly:

  * u_j for shot j: Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC 2011) block
    g = j // 4 with counter (g, 0, 0, 0) and key (seed, 0), word j % 4,
    u_j = (word >> 11) * 2^-53;
  * the state is cut into blocks of min(2^n, 4096) amplitudes; x_j = u_j * P
    with P the total of |psi|^2; the block is the largest b with
    prefix[b] <= x_j, moved down over zero-mass blocks; within the block
    (64 segments of 64, see _within_block) the first segment and then the
    first amplitude whose running sum exceeds x_j - prefix[b], moved down
    over zero probabilities.

How it is pinned (tests/test_sampling.py): the Philox restatement against
numpy's Philox bit generator (an independent implementation of the same
generator) and against the plain-integer rounds below; the sampled
distribution against the exact probabilities (chi-square, and exact answers
for states with one nonzero amplitude or zero-mass blocks).
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

_M0, _M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
_W0, _W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
_MASK = (1 << 64) - 1
BLOCK = 4096


def philox4x64_10(ctr: List[int], key: List[int]) -> List[int]:
    """Ten Philox4x64 rounds on Python integers (the published round:
    (hi, lo) = M * x for x0 and x2, outputs hi1^x1^k0, lo1, hi0^x3^k1, lo0,
    key bumped by the Weyl constants after each round)."""
    x0, x1, x2, x3 = (int(v) & _MASK for v in ctr)
    k0, k1 = (int(v) & _MASK for v in key)
    for _ in range(10):
        p0, p1 = _M0 * x0, _M1 * x2
        x0, x1, x2, x3 = (p1 >> 64) ^ x1 ^ k0, p1 & _MASK, (p0 >> 64) ^ x3 ^ k1, p0 & _MASK
        k0, k1 = (k0 + _W0) & _MASK, (k1 + _W1) & _MASK
    return [x0, x1, x2, x3]


def stream_words(seed: int, nwords: int) -> np.ndarray:
    """Words 0..nwords-1 of the stream: block g (counter (g, 0, 0, 0), key
    (seed, 0)), four words per block.  numpy's Philox increments its 256-bit
    counter before each block, so it starts from counter -1."""
    nblk = (int(nwords) + 3) // 4
    if nblk == 0:
        return np.zeros(0, dtype=np.uint64)
    bg = np.random.Philox(key=np.array([int(seed) & _MASK, 0], dtype=np.uint64),
                          counter=np.array([_MASK] * 4, dtype=np.uint64))
    return bg.random_raw(4 * nblk)[: int(nwords)].astype(np.uint64)


def uniforms(seed: int, shots: int) -> np.ndarray:
    w = stream_words(seed, shots)
    return (w >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def probabilities(psi: np.ndarray) -> np.ndarray:
    """|psi_i|^2 in float64 from the complex64 amplitudes (each square exact)."""
    a = np.asarray(psi)
    re = a.real.astype(np.float64)
    im = a.imag.astype(np.float64)
    return re * re + im * im


def sample_counts(psi: np.ndarray, shots: int, seed: int) -> Tuple[np.ndarray, np.ndarray]:
    """(index, count) of the distinct sampled states, index ascending."""
    p = probabilities(psi)
    dim = p.size
    blk = min(dim, BLOCK)
    nb = dim // blk
    bsum = p.reshape(nb, blk).sum(axis=1)
    prefix = np.concatenate([[0.0], np.cumsum(bsum)])
    total = prefix[-1]
    if not (total > 0.0 and np.isfinite(total)):
        raise ValueError("the state has zero or non-finite norm")
    x = uniforms(seed, shots) * total
    b = np.searchsorted(prefix[:nb], x, side="right") - 1
    for j in np.nonzero((b > 0) & (prefix[b + 1] == prefix[b]))[0]:
        while b[j] > 0 and prefix[b[j] + 1] == prefix[b[j]]:
            b[j] -= 1
    xl = x - prefix[b]
    out = np.empty(int(shots), dtype=np.int64)
    for blkid in np.unique(b):
        sel = np.nonzero(b == blkid)[0]
        pv = np.zeros(BLOCK)
        pv[:blk] = p[blkid * blk:(blkid + 1) * blk]
        k = _within_block(pv, xl[sel])
        out[sel] = blkid * blk + k
    idx, cnt = np.unique(out, return_counts=True)
    return idx.astype(np.uint64), cnt.astype(np.int64)


def _within_block(pv: np.ndarray, xl: np.ndarray) -> np.ndarray:
    """Amplitude of each shot inside a block of BLOCK probabilities: the block
    is 64 segments of 64; the segment is the first s with E[s] + S[s] > x
    (segment sums S, exclusive prefix E; the last one if none), the amplitude
    the first l with E[s] + cumsum(segment)[l] > x (lane 63 if none), moved
    down over zero probabilities."""
    seg_p = pv.reshape(64, 64)
    S = seg_p.sum(axis=1)
    E = np.concatenate([[0.0], np.cumsum(S)[:-1]])
    I = E + S
    hit = I[None, :] > xl[:, None]
    seg = np.where(hit.any(axis=1), hit.argmax(axis=1), 63)
    c = E[seg][:, None] + np.cumsum(seg_p[seg], axis=1)
    hit2 = c > xl[:, None]
    lane = np.where(hit2.any(axis=1), hit2.argmax(axis=1), 63)
    k = 64 * seg + lane
    for t in range(k.size):
        while k[t] > 0 and pv[k[t]] == 0.0:
            k[t] -= 1
    return k


def reference_counts(psi: np.ndarray, shots: int, seed: int) -> Tuple[np.ndarray, np.ndarray]:
    """The reference's host path (engine.py:377-465 without noise): |psi|^2,
    normalised, ``choice(dim, size=shots, p=p)``, ``bincount``.  Used as the
    timed CPU baseline of the GPU sampler (same distribution, numpy's stream)."""
    p = probabilities(psi)
    s = float(p.sum())
    p = p / s if s > 0 else np.full(p.size, 1.0 / p.size)
    rng = np.random.default_rng(seed)
    draws = rng.choice(p.size, size=int(shots), p=p)
    cnt = np.bincount(draws, minlength=p.size)
    idx = np.nonzero(cnt)[0]
    return idx.astype(np.uint64), cnt[idx].astype(np.int64)

