"""Batched engine on the host: the persistent kernel's per-thread code replayed
workgroup by workgroup (between guard bands) must reproduce the oracle for
every problem of a mixed batch, for several grid sizes."""
import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import data, workloads as W


def run(problems, grid, seed=0, device=None):
    ins = [data.uniform(int(np.prod(d)), e, seed + i) for i, (d, p, e) in enumerate(problems)]
    outs = L.batch_host_replay(problems, ins, grid=grid, device=device)
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


def test_multi_item_units():
    # A one-CU device gives a four-workgroup grid, so units grow to many
    # tiles / outer indices each, with ragged last units.
    dev = L.mi355x_desc()
    dev.cu_count = 1
    rng = np.random.default_rng(5)
    probs = [W.random_problem(rng, 2, 5, 20000, 200000, [4, 8, 16]) for _ in range(12)]
    probs += [([1000, 999], [1, 0], 4), ([3001, 77], [1, 0], 8), ([64, 33, 50], [0, 2, 1], 4),
              ([2, 8, 16, 32, 3], [3, 1, 0, 2, 4], 4), ([513, 700], [0, 1], 8)]
    run(probs, 0, device=dev)
    run(probs, 3, device=dev)


CANDIDATE_CASES = [
    # The named large cases scaled down, keeping their permutations.
    ([5, 97, 211, 33], [3, 1, 0, 2], 8),
    ([4, 4, 8, 32, 32, 16], [2, 5, 0, 4, 1, 3], 8),
    ([2, 8, 16, 32, 64, 8], [5, 3, 1, 0, 2, 4], 4),
    ([256, 64, 32], [0, 2, 1], 8),
    ([512, 384], [1, 0], 4),
    ([33, 7, 130], [2, 0, 1], 16),
    ([3, 3, 3, 2, 4], [1, 3, 0, 4, 2], 16),
]


@pytest.mark.parametrize("dim,perm,elem", CANDIDATE_CASES)
def test_every_batch_candidate(dim, perm, elem):
    # The autotuner runs candidates other than the model's choice, so each
    # one, in both descriptor forms and at several unit sizes, must reproduce
    # the oracle between guard bands and stay inside the kernel's LDS.
    x = data.uniform(int(np.prod(dim)), elem, 3)
    ref = oracle.transpose(dim, perm, elem, x).view(np.uint8)
    n = L.batch_host_candidates(dim, perm, elem)
    assert n >= 1
    for i in range(n):
        for fb in (False, True):
            for ub in (16 << 10, 160 << 10):
                r, y = L.batch_host_replay_candidate(dim, perm, elem, x, i, fallback=fb, unit_bytes=ub, grid=37)
                assert r == L.LIBRETT_SUCCESS, (i, fb, ub, r)
                assert np.array_equal(ref, y.view(np.uint8)), (i, fb, ub)


def test_batch_candidates_random():
    rng = np.random.default_rng(23)
    for _ in range(25):
        d, p, e = W.random_problem(rng, 3, 5, 1000, 60000, [8, 16])
        x = data.uniform(int(np.prod(d)), e, 4)
        ref = oracle.transpose(d, p, e, x).view(np.uint8)
        for i in range(L.batch_host_candidates(d, p, e)):
            r, y = L.batch_host_replay_candidate(d, p, e, x, i, unit_bytes=32 << 10)
            assert r == L.LIBRETT_SUCCESS and np.array_equal(ref, y.view(np.uint8)), (d, p, e, i)


def test_batch_covers_every_method():
    # Shapes that make the batch planner pick each kind of kernel.
    probs = [([256, 96], [1, 0], 8), ([64, 32, 16], [0, 2, 1], 4), ([2, 8, 16, 32], [3, 1, 0, 2], 4),
             ([5, 7, 9, 11], [2, 3, 0, 1], 16), ([1000], [0], 16), ([3, 3, 3, 2, 4], [1, 3, 0, 4, 2], 8)]
    run(probs, 0)
    run(probs, 3)


def test_batch_rejects_bad_problem():
    with pytest.raises(ValueError):
        L.batch_host_replay([([4, 4], [0, 0], 8)], [np.zeros(16)])
