"""The float32 kernels' dispatch order of a posterior batch (CPU): a
permutation of the chains, each XCD range filled to its capacity, the costliest
posteriors first inside a range, loads balanced across the 8 ranges."""
import random

import torch

from geomjax_amd.models import _posterior_table


class _Post:
    def __init__(self, N, D):
        self.N, self.D = N, D
        self.X = torch.zeros(N, D, dtype=torch.float64)
        self.y = torch.zeros(N, dtype=torch.float64)


def _cost(m):
    w = 32 if m.D <= 32 else (64 if m.D <= 64 else 128)
    return m.N * w * w


def test_dispatch_order_is_a_balanced_permutation():
    rng = random.Random(0)
    for trial in range(5):
        ms = [_Post(rng.choice([256, 1000, 4096]), rng.choice([8, 16, 32, 64, 128])) for _ in range(rng.randint(2, 60))]
        counts = [rng.randint(1, 40) for _ in ms]
        cp = torch.repeat_interleave(torch.arange(len(ms), dtype=torch.int32), torch.tensor(counts))
        C = int(sum(counts))
        order = _posterior_table(ms, cp, "cpu").dispatch_order(cp, C).tolist()
        assert sorted(order) == list(range(C))
        per = (C + 7) // 8
        loads = []
        for x in range(8):
            rng_x = order[x * per:(x + 1) * per]
            assert len(rng_x) == max(0, min(per, C - x * per))
            costs = [_cost(ms[int(cp[c])]) for c in rng_x]
            assert costs == sorted(costs, reverse=True)  # costliest first
            loads.append(sum(costs))
        # no range carries more than the mean plus one chain of the costliest posterior per posterior split
        assert max(loads) <= sum(loads) / 8 + max(_cost(m) * n for m, n in zip(ms, counts))


def test_single_posterior_needs_no_order():
    ms = [_Post(100, 8)]
    cp = torch.zeros(16, dtype=torch.int32)
    assert _posterior_table(ms, cp, "cpu").dispatch_order(cp, 16) is None
