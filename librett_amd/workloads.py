"""Seeded random transpose problems for benchmarks, tuning and training.

Shapes follow the published workload descriptions: a rank, dims whose
product lands near a target volume, and a random non-identity permutation.
Classes for the large single transposes:
  keep  : perm[0] == 0 and dim0 in [64, 4096]   (TiledCopy territory)
  move  : perm[0] != 0, dim0 and dim[perm[0]] >= 64 (Tiled territory)
  small : dim0 in [2, 8], perm[0] != 0          (Packed / PackedSplit)
"""
import math

import numpy as np


def _perm(rng, rank, first=None, avoid_identity=True):
    while True:
        if first is None:
            p = [int(x) for x in rng.permutation(rank)]
        else:
            rest = [int(x) for x in rng.permutation([i for i in range(rank) if i != first])]
            p = [first] + rest
        if not avoid_identity or p != list(range(rank)) or rank == 1:
            return p


def _fill_dims(rng, rank, fixed, target):
    """Dims (log-uniform) for positions not in `fixed` so the volume is near target."""
    free = [i for i in range(rank) if i not in fixed]
    dims = [0] * rank
    rest = target
    for i, v in fixed.items():
        dims[i] = v
        rest /= v
    rest = max(rest, 2.0 ** len(free))
    if free:
        w = rng.uniform(0.3, 1.0, size=len(free))
        w = w / w.sum()
        for i, wi in zip(free, w):
            dims[i] = max(2, int(round(rest ** wi)))
        # correct the largest free dim toward the target volume
        vol = math.prod(dims)
        j = max(free, key=lambda k: dims[k])
        dims[j] = max(2, int(round(dims[j] * target / vol)))
    return dims


def large_case(seed, elem, cls, target_bytes=256 << 20):
    """One ~256 MiB single-transpose problem of class keep/move/small."""
    rng = np.random.default_rng(seed)
    target = target_bytes // elem
    rank = int(rng.integers(2, 7))
    if cls == "keep":
        rank = max(rank, 3)
        d0 = int(2 ** rng.uniform(6, 12))
        dims = _fill_dims(rng, rank, {0: d0}, target)
        perm = _perm(rng, rank, first=0)
    elif cls == "move":
        while True:
            perm = _perm(rng, rank)
            if perm[0] != 0:
                break
        d0 = int(2 ** rng.uniform(6, 12))
        if rank == 2:
            dims = [d0, max(64, int(round(target / d0)))]
        else:
            dp = int(2 ** rng.uniform(6, 12))
            dims = _fill_dims(rng, rank, {0: d0, perm[0]: dp}, target)
    else:
        rank = max(rank, 3)
        while True:
            perm = _perm(rng, rank)
            if perm[0] != 0:
                break
        dims = _fill_dims(rng, rank, {0: int(rng.integers(2, 9))}, target)
    return dims, perm


def random_problem(rng, rank_lo, rank_hi, vol_lo, vol_hi, elems):
    """Rank U{lo..hi}, volume log-uniform in [vol_lo, vol_hi] elements,
    element size drawn from `elems`, random non-identity permutation."""
    rank = int(rng.integers(rank_lo, rank_hi + 1))
    elem = int(rng.choice(elems))
    target = math.exp(rng.uniform(math.log(vol_lo), math.log(vol_hi)))
    dims = _fill_dims(rng, rank, {}, target)
    perm = _perm(rng, rank)
    return dims, perm, elem


def test2_dims(rng, rank, volmax=2e8, min_dim=2.0):
    """Dims drawn the way the reference's test2 draws them
    (tests/librett_test.cpp:194-259), from a seeded generator instead of
    rand(): a target volume uniform in [(min_dim+1)^rank, volmax], dims
    spread around its rank-th root, rescaled toward the target and kept >= 2,
    redrawn until the volume fits volmax and is within a factor 2.3 of the
    target. The reference carries the running volume from one rescaling
    attempt into the next, and so does this restatement."""
    volmin = (min_dim + 1.0) ** rank
    dims = [2] * rank
    curvol = 1.0
    for _ in range(1000):
        vol = volmin + (volmax - volmin) * rng.random()
        for _ in range(50):
            ave = vol ** (1.0 / rank)
            spread = ave - min_dim
            for r in range(rank):
                rn = 2.0 * (rng.random() - 0.5)
                dims[r] = int(ave + spread * rn)
                curvol *= dims[r]
            scale = (vol / curvol) ** (1.0 / rank)
            curvol = 1.0
            for r in range(rank):
                dims[r] = max(2, int(round(dims[r] * scale)))
                curvol *= dims[r]
            if not (curvol > volmax or abs(curvol - vol) / vol > 2.3):
                break
        if not (curvol > volmax or abs(curvol - vol) / vol > 2.3):
            return list(dims)
    raise RuntimeError("no dims found for rank %d" % rank)


def batched_problems(n=4096, seed=20261202):
    """The batched workload: `n` problems of rank 3-5, volume log-uniform in
    [1e3, 1e6] elements, float64 or complex128, random non-identity
    permutations."""
    rng = np.random.default_rng(seed)
    return [random_problem(rng, 3, 5, 1e3, 1e6, [8, 16]) for _ in range(n)]
