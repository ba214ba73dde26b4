"""Host-side planner and kernel index maps, checked without a GPU.

* rank reduction (reference src/plan.cpp:64-127) by brute force;
* every candidate plan of a problem, executed by the host replay of the
  kernels' per-thread code (same source as the device kernels), must equal
  the oracle bit for bit. Shapes follow the reference's test1 (all
  permutations, dims 2 + r*14/rank, tests/librett_test.cpp:168-188), test3
  (hand-picked, :277-417, scaled where the host replay would be slow) and the
  bench7 rank-8/12 shapes (tests/librett_bench.cpp:718,738, scaled).
"""
import itertools

import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import data

RNG = np.random.default_rng(1202)


def replay_all(dim, perm, elem, seed=1, device=None, max_cands=None):
    vol = int(np.prod(dim))
    x = data.uniform(vol, elem, seed)
    ref = oracle.transpose(dim, perm, elem, x)
    cands = L.host_candidates(dim, perm, elem, device=device)
    assert cands, (dim, perm, elem)
    idx = range(len(cands)) if max_cands is None else range(min(max_cands, len(cands)))
    for i in idx:
        y = L.host_replay(dim, perm, elem, i, x, device=device)
        assert np.array_equal(y.view(np.uint8), ref.view(np.uint8)), (dim, perm, elem, cands[i]["desc"])
    return cands


def test_reduce_ranks_examples():
    assert L.reduce_ranks([64, 48, 32], [2, 0, 1]) == ([3072, 32], [1, 0])
    assert L.reduce_ranks([2, 3, 4], [0, 1, 2]) == ([24], [0])
    # The size-1 dim is dropped first, leaving (5, 7) with perm (1, 0).
    assert L.reduce_ranks([5, 1, 7], [2, 1, 0]) == ([5, 7], [1, 0])
    d, p = L.reduce_ranks([5, 1, 7], [2, 1, 0])
    assert int(np.prod(d)) == 35 and sorted(p) == list(range(len(p)))


def test_reduce_ranks_brute_force():
    for rank in range(1, 7):
        for perm in itertools.permutations(range(rank)):
            dim = [int(v) for v in RNG.integers(2, 5, size=rank)]
            rd, rp = L.reduce_ranks(dim, list(perm))
            assert int(np.prod(rd)) == int(np.prod(dim))
            assert sorted(rp) == list(range(len(rp)))
            # no fusable run remains
            assert all(rp[i] != rp[i - 1] + 1 for i in range(1, len(rp)))
            # the reduced problem is the same map
            x = data.uniform(int(np.prod(dim)), 4, rank)
            a = oracle.transpose(dim, list(perm), 4, x)
            b = oracle.transpose(rd, rp, 4, x)
            assert np.array_equal(a, b)


@pytest.mark.parametrize("elem", [4, 8, 16])
def test_all_perms_rank2_to_5(elem):
    for rank in range(2, 6):
        dim = [2 + r * (16 - 2) // rank for r in range(rank)]
        for perm in itertools.permutations(range(rank)):
            replay_all(dim, list(perm), elem, seed=rank)


@pytest.mark.parametrize("elem", [4, 8])
def test_rank6_sampled_perms(elem):
    rank = 6
    dim = [2 + r * (16 - 2) // rank for r in range(rank)]
    perms = list(itertools.permutations(range(rank)))
    for k in RNG.choice(len(perms), size=24, replace=False):
        replay_all(dim, list(perms[k]), elem, seed=int(k))


@pytest.mark.parametrize("elem", [4, 8, 16])
def test_hand_picked_shapes(elem):
    cases = [
        ([43, 67], [1, 0]),
        ([4096, 2], [1, 0]),
        ([65, 59, 22], [0, 2, 1]),
        ([24, 33, 16, 7], [1, 0, 2, 3]),
        ([2, 5, 9, 12], [0, 1, 2, 3]),
        ([2, 4, 6, 9, 11, 13], [0, 1, 2, 3, 4, 5]),
        ([5, 12, 15, 6, 7], [1, 3, 4, 2, 0]),
        ([5, 3, 2, 9, 14], [0, 1, 3, 2, 4]),
        ([3, 3, 3, 2, 4], [1, 3, 0, 4, 2]),
        ([24, 32, 16, 3, 4, 9], [5, 1, 4, 2, 3, 0]),
        ([130, 70], [1, 0]),
        ([1, 7, 1, 9], [3, 2, 1, 0]),
        ([1, 1, 1], [2, 0, 1]),
    ]
    for dim, perm in cases:
        replay_all(dim, perm, elem)


def test_f32_eight_byte_lanes():
    # Plane extents that are even but not multiples of 4 get 8-byte lanes on
    # that side (4 and 2, 2 and 4, 2 and 2, 2 and 1, 1 and 2), full and
    # partial tiles.
    seen = set()
    for dim, perm in (([34, 18], [1, 0]), ([68, 50, 3], [1, 0, 2]), ([30, 36], [1, 0]), ([130, 46], [1, 0]),
                      ([6, 7, 10], [2, 1, 0]), ([9, 70], [1, 0]), ([70, 9], [1, 0])):
        for c in replay_all(dim, perm, 4):
            if c["method"] == 4:
                seen.add((c["vec_x"], c["vec_y"]))
    assert {(2, 2), (4, 2), (2, 4), (2, 1), (1, 2)} <= seen


def test_output_span_planes():
    # Planes whose y group is the output's whole leading run that holds no
    # x-group dim: y = 4 dims for (288,2,2,16,589,6) (the output's leading
    # dims 589, 2, 2, 6 before input dim 0), y = 2 dims with x = 2 dims for
    # (5,44,6,6,207,41); each replays bit-exact, scaled down.
    for dim, perm, want in (([12, 2, 2, 5, 33, 6], [4, 2, 1, 5, 0, 3], "groups 1/4"),
                            ([5, 14, 6, 6, 9, 3], [4, 2, 1, 5, 0, 3], "groups 2/2"),
                            ([11, 13, 7], [2, 1, 0], "groups 1/2")):
        for elem in (4, 8):
            descs = [c["desc"] for c in replay_all(dim, perm, elem)]
            assert any(want in d for d in descs), (dim, elem, want)


def expected_planes(dim, perm):
    """(x dims, y dims) of the Tiled planes the planner builds for a
    reduced problem (planner.cpp add_tiled): the reference's plane (1, 1);
    x the leading input dims reaching 64 elements and y the same on the
    output side, where dim 0 (or output dim 0) is short; and y the output's
    whole leading run that holds no x dim. Groups never share a dim."""
    n = len(dim)

    def reach(inp):
        v, k = 1, 0
        while k < n and k < 6:
            v *= dim[k if inp else perm[k]]
            k += 1
            if v >= 64:
                break
        return k

    def span(a):
        v, k = 1, 0
        while k < n and k < 6:
            d = perm[k]
            if d < a or v * dim[d] > (1 << 30):
                break
            v *= dim[d]
            k += 1
        return k

    a64, b64 = reach(True), reach(False)
    combos = [(a64, 1), (1, b64), (a64, b64), (1, span(1)), (a64, span(a64))]
    out = {(1, 1)}
    for i, (a, b) in enumerate(combos):
        if a < 1 or b < 1 or (a, b) == (1, 1):
            continue
        if i < 3 and ((a > 1 and dim[0] >= 64) or (b > 1 and dim[perm[0]] >= 64)):
            continue
        if set(range(a)) & set(perm[:b]):
            continue
        out.add((a, b))
    return out


def test_tiled_planes_follow_the_rules():
    # Every plane the rules name appears among the candidates (some tile
    # shape of it), and no other, on random problems at 4 and 8 bytes.
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(300):
        r = int(rng.integers(2, 7))
        dim = [int(v) for v in rng.integers(2, 40, r)]
        perm = [int(v) for v in rng.permutation(r)]
        rd, rp = L.reduce_ranks(dim, perm)
        if len(rd) < 2 or rp[0] == 0:
            continue
        want = expected_planes(rd, rp)
        for elem in (4, 8):
            got = set()
            for c in L.host_candidates(dim, perm, elem):
                if c["method"] != 4:
                    continue
                d = c["desc"]
                if " groups " in d:
                    a, b = d.split(" groups ")[1].split(" ")[0].split("/")
                    got.add((int(a), int(b)))
                else:
                    got.add((1, 1))
            assert got == want, (dim, perm, rd, rp, elem, got, want)
            checked += 1
    assert checked > 200


def test_bench7_high_rank_scaled():
    # tests/librett_bench.cpp:718,738 shapes scaled down for the host replay.
    replay_all([5, 3, 2, 4, 3, 3, 3, 4], [7, 2, 5, 0, 1, 3, 6, 4], 8)
    replay_all([2, 3, 4, 3, 2, 2, 3, 2, 2, 2, 2, 3], [11, 0, 5, 3, 9, 1, 10, 2, 6, 4, 8, 7], 4)


def test_random_shapes():
    for _ in range(60):
        rank = int(RNG.integers(2, 7))
        dim = [int(d) for d in RNG.integers(1, 9, size=rank)]
        perm = [int(p) for p in RNG.permutation(rank)]
        elem = int(RNG.choice([4, 8, 16]))
        replay_all(dim, perm, elem, seed=int(RNG.integers(1 << 20)))


def test_small_lds_device_forces_split():
    # A device with 4 KiB of LDS forces PackedSplit candidates on shapes that
    # would otherwise stage whole; they must still be exact.
    dev = L.DeviceDesc(16, 4096, 4096, 1024, 32, 1e12)
    cands = replay_all([40, 37, 6], [2, 0, 1], 8, device=dev)
    assert any(c["method"] == L.PACKED_SPLIT for c in cands)
    replay_all([9, 70, 33], [1, 2, 0], 4, device=dev)


def test_candidate_methods_follow_reference_rules():
    # Reduced rank 1 -> Trivial only (src/plan.cpp:708-710).
    c = L.host_candidates([4, 5, 6], [0, 1, 2], 8)
    assert [x["method"] for x in c] == [L.TRIVIAL]
    # perm[0] == 0 after reduction -> TiledCopy, never Tiled.
    m = {x["method"] for x in L.host_candidates([64, 32, 16], [0, 2, 1], 4)}
    assert L.TILED_COPY in m and L.TILED not in m
    # perm[0] != 0 -> Tiled, never TiledCopy.
    m = {x["method"] for x in L.host_candidates([64, 32, 16], [2, 0, 1], 4)}
    assert L.TILED in m and L.TILED_COPY not in m
    # Small leading dim: Packed candidates exist.
    m = {x["method"] for x in L.host_candidates([2, 8, 16, 32, 64, 128], [5, 3, 1, 0, 2, 4], 4)}
    assert L.PACKED in m or L.PACKED_SPLIT in m


def test_model_prefers_tiled_for_large_square():
    c = L.host_candidates([8192, 8192], [1, 0], 4)
    best = min(c, key=lambda x: x["predicted_s"])
    assert best["method"] == L.TILED
    assert best["bytes"] == 2 * 8192 * 8192 * 4
