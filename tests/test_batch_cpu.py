"""Batched engine on the host: the persistent kernel's per-thread code replayed
workgroup by workgroup (between guard bands) must reproduce the oracle for
every problem of a mixed batch, for several grid sizes."""
import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import data, workloads as W


def run(problems, grid, seed=0):
    ins = [data.uniform(int(np.prod(d)), e, seed + i) for i, (d, p, e) in enumerate(problems)]
    outs = L.batch_host_replay(problems, ins, grid=grid)
    for (d, p, e), x, y in zip(problems, ins, outs):
        ref = oracle.transpose(d, p, e, x)
        assert np.array_equal(ref.view(np.uint8), y.view(np.uint8)), (d, p, e)


@pytest.mark.parametrize("grid", [0, 1, 7, 300])
def test_mixed_batch(grid):
    rng = np.random.default_rng(11)
    probs = [W.random_problem(rng, 3, 5, 100, 30000, [4, 8, 16]) for _ in range(60)]
    probs = [(d, p, e) for d, p, e in probs]
    probs += [([3, 4, 5], [0, 1, 2], 8), ([37, 1, 5], [2, 1, 0], 16), ([7], [0], 4), ([64, 48, 32], [2, 0, 1], 8)]
    run(probs, grid)


def test_batch_covers_every_method():
    # Shapes that make the batch planner pick each kind of kernel.
    probs = [([256, 96], [1, 0], 8), ([64, 32, 16], [0, 2, 1], 4), ([2, 8, 16, 32], [3, 1, 0, 2], 4),
             ([5, 7, 9, 11], [2, 3, 0, 1], 16), ([1000], [0], 16), ([3, 3, 3, 2, 4], [1, 3, 0, 4, 2], 8)]
    run(probs, 0)
    run(probs, 3)


def test_batch_rejects_bad_problem():
    with pytest.raises(ValueError):
        L.batch_host_replay([([4, 4], [0, 0], 8)], [np.zeros(16)])
