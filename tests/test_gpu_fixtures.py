"""The reference's own test and benchmark inputs on the GPU, at full size.

- test1 (tests/librett_test.cpp:168-188): every permutation of rank 7 at 4
  and 8 bytes (ranks 2-6 are in test_gpu_parity.py), plus every permutation
  of ranks 2-5 at 16 bytes, which the reference never instantiated.
- test2 (tests/librett_test.cpp:194-272): ranks 8-15, dims drawn the way the
  reference draws them (seeded here), volumes up to 2e8 elements, 4 and 8
  bytes; and small problems of the same ranks with every candidate plan.
- bench6, the TTC suite (tests/librett_bench.cpp:576-708): 57 (dims, perm)
  pairs at 8 bytes.
- bench7 (tests/librett_bench.cpp:713-757): the rank-8 and rank-12 shapes at
  full size with the reversed, identity and shuffled permutations.

Outputs too large for the host oracle are checked on the device with the
index-pattern checker (every element compared with the input element it
must come from); small ones are compared with the oracle bit for bit.
"""
import itertools
import json
import os

import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import data, hip, workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = json.load(open(os.path.join(HERE, "golden", "ttc_cases.json")))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def huge():
    # bench7's rank-12 shape at 8 bytes is 2.63 GB.
    n = 328458240 * 8 + 256
    a, b = hip.DeviceBuffer(n), hip.DeviceBuffer(n)
    yield a, b
    a.free()
    b.free()


def check_on_device(dim, perm, elem, stream, bufs, fill=True):
    a, b = bufs
    n = int(np.prod(dim)) * elem
    if fill:
        L.check_fill(a.ptr, n, stream)
    hip.memset(b.ptr, 0xEE, n)
    r, h = L.librettPlan(len(dim), list(dim), list(perm), elem, stream)
    assert r == L.LIBRETT_SUCCESS, (dim, perm, elem, r)
    desc = L.plan_info(h)["desc"]
    assert L.librettExecute(h, a.ptr, b.ptr) == L.LIBRETT_SUCCESS
    bad, first = L.check_transpose(b.ptr, list(dim), list(perm), elem, stream)
    assert L.librettDestroy(h) == L.LIBRETT_SUCCESS
    assert bad == 0, (dim, perm, elem, desc, bad, first)
    return desc


def check_oracle(dim, perm, elem, stream, a, b, x, ref, index=None):
    hip.htod(a.ptr, x)
    hip.memset(b.ptr, 0xEE, x.nbytes)
    if index is None:
        r, h = L.librettPlan(len(dim), dim, perm, elem, stream)
    else:
        r, h = L.plan_candidate(len(dim), dim, perm, elem, stream, index)
    assert r == L.LIBRETT_SUCCESS, (dim, perm, elem, index, r)
    desc = L.plan_info(h)["desc"]
    assert L.librettExecute(h, a.ptr, b.ptr) == L.LIBRETT_SUCCESS
    hip.stream_sync(stream)
    assert L.librettDestroy(h) == L.LIBRETT_SUCCESS
    out = np.empty_like(x)
    hip.dtoh(out, b.ptr)
    assert np.array_equal(out.view(np.uint8), ref.view(np.uint8)), (dim, perm, elem, desc)


@pytest.mark.parametrize("elem", [4, 8])
def test_test1_rank7_every_permutation(gpu_stream, huge, elem):
    rank = 7
    dim = [2 + r * (16 - 2) // rank for r in range(rank)]
    say("case: test1 rank 7, all 5040 permutations, elem", elem)
    L.check_fill(huge[0].ptr, int(np.prod(dim)) * elem, gpu_stream)
    for k, perm in enumerate(itertools.permutations(range(rank))):
        if k % 1000 == 0:
            say("   perm", k)
        check_on_device(dim, perm, elem, gpu_stream, huge, fill=False)


def test_all_perms_16_bytes(gpu_stream):
    a, b = hip.DeviceBuffer(1 << 20), hip.DeviceBuffer(1 << 20)
    try:
        for rank in range(2, 6):
            dim = [2 + r * (16 - 2) // rank for r in range(rank)]
            x = data.uniform(int(np.prod(dim)), 16, rank)
            say("case: all perms rank", rank, dim, "16 B")
            for perm in itertools.permutations(range(rank)):
                perm = list(perm)
                check_oracle(dim, perm, 16, gpu_stream, a, b, x, oracle.transpose(dim, perm, 16, x))
    finally:
        a.free()
        b.free()


@pytest.mark.parametrize("rank", list(range(8, 16)))
def test_test2_style_large(gpu_stream, huge, rank):
    rng = np.random.default_rng(20261202 + rank)
    for k in range(6):
        dim = W.test2_dims(rng, rank)
        perm = [int(v) for v in rng.permutation(rank)]
        for elem in (8, 4):
            say("case: test2 rank", rank, "sample", k, dim, perm, elem, "vol %d" % int(np.prod(dim)))
            desc = check_on_device(dim, perm, elem, gpu_stream, huge)
            say("   ->", desc)


@pytest.mark.parametrize("rank", list(range(8, 16)))
def test_test2_style_every_candidate(gpu_stream, rank):
    # Small volumes so every candidate plan can be run and compared with the
    # oracle; the planner's staged-dim limit (12) and its outer map are used
    # at these ranks.
    rng = np.random.default_rng(7 + rank)
    dim = W.test2_dims(rng, rank, volmax=4e4)
    perm = [int(v) for v in rng.permutation(rank)]
    a, b = hip.DeviceBuffer(1 << 20), hip.DeviceBuffer(1 << 20)
    try:
        for elem in (4, 8, 16):
            x = data.uniform(int(np.prod(dim)), elem, rank)
            ref = oracle.transpose(dim, perm, elem, x)
            n = len(L.candidates(dim, perm, elem))
            say("case: test2 every candidate rank", rank, dim, perm, elem, "n =", n)
            for i in range(n):
                check_oracle(dim, perm, elem, gpu_stream, a, b, x, ref, index=i)
    finally:
        a.free()
        b.free()


def test_ttc_suite(gpu_stream, huge):
    cases = FIXTURES["ttc"]
    assert len(cases) == 57
    for k, c in enumerate(cases):
        say("case: TTC %d %s %s 8 B" % (k, c["dim"], c["perm"]))
        desc = check_on_device(c["dim"], c["perm"], 8, gpu_stream, huge)
        say("   ->", desc)


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("elem", [4, 8])
def test_bench7_shapes(gpu_stream, huge, which, elem):
    base = FIXTURES["bench7"][which]
    rank = len(base)
    rng = np.random.default_rng(717 + which)
    cases = [(base, list(range(rank - 1, -1, -1))), (base, list(range(rank)))]
    for _ in range(4):
        d = [int(v) for v in rng.permutation(base)]
        p = [int(v) for v in rng.permutation(rank)]
        if p != list(range(rank)):
            cases.append((d, p))
    for dim, perm in cases:
        say("case: bench7 rank %d %s %s %d B (vol %d)" % (rank, dim, perm, elem, int(np.prod(dim))))
        desc = check_on_device(dim, perm, elem, gpu_stream, huge)
        say("   ->", desc)
