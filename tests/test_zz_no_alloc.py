"""tqsv_dot_noalloc (GPU): the allocation-free variant of tqsv_dot gives the
same sums as tqsv_dot through the C ABI, on two calls in a row (the second on
a smaller state, so that its partial sums only partly overwrite the fixed
scratch array of the first), for several flip and sign masks."""
from __future__ import annotations

import numpy as np
import pytest

from tests.gpu_util import rand_state, to_gpu

pytestmark = pytest.mark.gpu


def test_dot_noalloc_matches_dot_twice_second_smaller(gpu):
    import torch

    from tyxonq_amd import _native as nat

    rng = np.random.default_rng(77)
    out = torch.full((4,), float("nan"), dtype=torch.float64, device=gpu)
    for n, masks in ((22, [(0, 0), (0b1010, 0b0110), ((1 << 21) | 5, 1 << 20)]),
                     (9, [(0, 0), (0b101, 0b011), (1 << 8, (1 << 8) | 1)])):
        u = to_gpu(rand_state(rng, n))
        v = to_gpu(rand_state(rng, n))
        for x, z in masks:
            want = nat.dot(u, v, x, z)
            nat.dot_noalloc(u, v, out[1:3], x, z)
            torch.cuda.synchronize()
            got = out[1:3].cpu().numpy()
            assert got[0] == want.real and got[1] == want.imag, (n, x, z, got, want)
        assert np.isnan(out[0].item()) and np.isnan(out[3].item())  # nothing written outside


def test_dot_noalloc_rejects_bad_arguments(gpu):
    import torch

    from tyxonq_amd import _native as nat

    u = to_gpu(rand_state(np.random.default_rng(1), 6))
    with pytest.raises(ValueError):
        nat.dot_noalloc(u, u, torch.zeros(1, dtype=torch.float64, device=gpu))
    with pytest.raises(ValueError):
        nat.dot_noalloc(u, u, torch.zeros(2, dtype=torch.float64, device=gpu), 1 << 6, 0)
