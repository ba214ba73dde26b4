"""Parity of the HIP kernels with the CPU oracle, through the C-ABI.

Every comparison is bit-exact on the elements' bytes. Sizes stay where the
oracle finishes in seconds, except the large named cases, which the OpenMP
oracle also handles in seconds.
"""
import itertools
import os
import subprocess
import sys

import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import data, hip

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)


class Bufs:
    """A pair of device buffers reused across cases."""

    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.a = hip.DeviceBuffer(nbytes + 256)
        self.b = hip.DeviceBuffer(nbytes + 256)

    def free(self):
        self.a.free()
        self.b.free()


@pytest.fixture(scope="module")
def bufs():
    b = Bufs(64 << 20)
    yield b
    b.free()


@pytest.fixture(scope="module")
def big_bufs():
    b = Bufs(600 << 20)
    yield b
    b.free()


def run_device(dim, perm, elem, x, stream, bufs, index=None, measure=False, offset=(0, 0)):
    nbytes = x.nbytes
    assert nbytes + 256 <= bufs.a.nbytes
    pin = bufs.a.ptr + offset[0]
    pout = bufs.b.ptr + offset[1]
    hip.htod(pin, x)
    hip.memset(pout, 0xEE, nbytes)
    if measure:
        r, h = L.librettPlanMeasure(len(dim), dim, perm, elem, stream, pin, pout)
    elif index is None:
        r, h = L.librettPlan(len(dim), dim, perm, elem, stream)
    else:
        r, h = L.plan_candidate(len(dim), dim, perm, elem, stream, index)
    assert r == L.LIBRETT_SUCCESS, (dim, perm, elem, r)
    info = L.plan_info(h)
    assert L.librettExecute(h, pin, pout) == L.LIBRETT_SUCCESS
    hip.stream_sync(stream)
    assert L.librettDestroy(h) == L.LIBRETT_SUCCESS
    out = np.empty_like(x)
    hip.dtoh(out, pout)
    return out, info


def check(dim, perm, elem, stream, bufs, seed=1, **kw):
    vol = int(np.prod(dim))
    x = data.uniform(vol, elem, seed)
    ref = oracle.transpose(dim, perm, elem, x, variant="omp")
    out, info = run_device(dim, perm, elem, x, stream, bufs, **kw)
    assert np.array_equal(out.view(np.uint8), ref.view(np.uint8)), (dim, perm, elem, info["desc"])
    return info


def test_config1_plain(gpu_stream, bufs):
    say("case: config 1 (64,48,32) f64 perm (2,0,1)")
    info = check([64, 48, 32], [2, 0, 1], 8, gpu_stream, bufs, seed=20261202)
    assert info["method"] in (L.TILED, L.PACKED, L.PACKED_SPLIT)


@pytest.mark.parametrize("elem", [4, 8, 16])
def test_example_cpp_case(gpu_stream, bufs, elem):
    say("case: example.cpp (3,3,3,2,4) perm (1,3,0,4,2) elem", elem)
    check([3, 3, 3, 2, 4], [1, 3, 0, 4, 2], elem, gpu_stream, bufs)


@pytest.mark.parametrize("elem", [4, 8, 16])
def test_every_candidate_small_shapes(gpu_stream, bufs, elem):
    shapes = [
        ([64, 48, 32], [2, 0, 1]),
        ([43, 67], [1, 0]),
        ([130, 70, 3], [1, 0, 2]),
        ([65, 59, 22], [0, 2, 1]),
        ([5, 12, 15, 6, 7], [1, 3, 4, 2, 0]),
        ([2, 8, 16, 32, 6], [4, 3, 1, 0, 2]),
        ([24, 32, 16, 3, 4, 9], [5, 1, 4, 2, 3, 0]),
        ([7, 1, 300], [2, 1, 0]),
        ([34, 18, 5], [1, 0, 2]),   # f32: 8-byte lanes both ways
        ([130, 46], [1, 0]),        # f32: 8-byte lanes, partial tiles
        ([68, 50, 3], [1, 0, 2]),   # f32: 16-byte input lanes, 8-byte output lanes
        ([12, 2, 2, 5, 33, 6], [4, 2, 1, 5, 0, 3]),  # planes with y over the output's leading run
        ([5, 14, 6, 6, 9, 3], [4, 2, 1, 5, 0, 3]),
    ]
    for dim, perm in shapes:
        n = len(L.candidates(dim, perm, elem))
        say("case: all candidates", dim, perm, elem, "n =", n)
        for i in range(n):
            check(dim, perm, elem, gpu_stream, bufs, index=i, seed=i)


@pytest.mark.parametrize("elem", [4, 8])
def test_all_perms_test1_style(gpu_stream, bufs, elem):
    # tests/librett_test.cpp:168-188, ranks 2..6 with dims 2 + r*14/rank.
    for rank in range(2, 7):
        dim = [2 + r * (16 - 2) // rank for r in range(rank)]
        say("case: test1 rank", rank, "elem", elem)
        for perm in itertools.permutations(range(rank)):
            check(dim, list(perm), elem, gpu_stream, bufs, seed=rank)


def test_all_perms_rank7_elem8(gpu_stream, bufs):
    rank = 7
    dim = [2 + r * (16 - 2) // rank for r in range(rank)]
    perms = list(itertools.permutations(range(rank)))
    pick = np.random.default_rng(7).choice(len(perms), size=400, replace=False)
    say("case: test1 rank 7 (400 of 5040 perms)")
    vol = int(np.prod(dim))
    x = data.uniform(vol, 8, 7)
    for k in pick:
        perm = perms[k]
        ref = oracle.transpose(dim, list(perm), 8, x)
        out, info = run_device(dim, list(perm), 8, x, gpu_stream, bufs)
        assert np.array_equal(out.view(np.uint8), ref.view(np.uint8)), (perm, info["desc"])


def test_hand_picked_test3(gpu_stream, big_bufs):
    # tests/librett_test.cpp:277-417 at full size.
    cases = [
        ([43, 67], [1, 0]),
        ([65536 * 32, 2], [1, 0]),
        ([1305, 599, 88], [0, 2, 1]),
        ([24, 330, 64, 147], [1, 0, 2, 3]),
        ([2, 5, 9, 12], [0, 1, 2, 3]),
        ([2, 4, 6, 9, 11, 13], [0, 1, 2, 3, 4, 5]),
        ([5, 42, 75, 86, 57], [1, 3, 4, 2, 0]),
        ([5, 3, 2, 9, 14], [0, 1, 3, 2, 4]),
    ]
    for dim, perm in cases:
        for elem in (4, 8):
            say("case: test3", dim, perm, elem)
            check(dim, perm, elem, gpu_stream, big_bufs)


def test_plan_measure(gpu_stream, bufs):
    for dim, perm, elem in [([64, 48, 32], [2, 0, 1], 8), ([2, 8, 16, 32, 64], [4, 3, 1, 0, 2], 4),
                            ([256, 64, 64], [0, 2, 1], 16)]:
        say("case: measure", dim, perm, elem)
        info = check(dim, perm, elem, gpu_stream, bufs, measure=True)
        assert info["measured_s"] > 0


def test_unaligned_pointers_use_fallback(gpu_stream, bufs):
    # Offsets of one element defeat 16-byte vector accesses.
    for dim, perm, elem in [([256, 128], [1, 0], 4), ([64, 32, 8], [0, 2, 1], 4), ([128, 96], [1, 0], 8),
                            ([130, 46], [1, 0], 4)]:
        for off in [(elem, 0), (0, elem), (elem, elem)]:
            say("case: unaligned", dim, perm, elem, off)
            check(dim, perm, elem, gpu_stream, bufs, offset=off)


def test_ten_streams_same_output():
    # tests/librett_test.cpp:422-488: ten plans on ten streams into one output
    # (171M float64 elements, 1.37 GB per tensor).
    dim, perm = [24, 32, 16, 36, 43, 9], [5, 1, 4, 2, 3, 0]
    vol = int(np.prod(dim))
    say("case: ten streams", dim, perm)
    x = data.uniform(vol, 8, 44)
    ref = oracle.transpose(dim, perm, 8, x, variant="omp")
    bufs = Bufs(x.nbytes)
    hip.htod(bufs.a.ptr, x)
    streams = [hip.stream_create() for _ in range(10)]
    handles = []
    for s in streams:
        r, h = L.librettPlan(len(dim), dim, perm, 8, s)
        assert r == L.LIBRETT_SUCCESS
        assert L.librettExecute(h, bufs.a.ptr, bufs.b.ptr) == L.LIBRETT_SUCCESS
        handles.append(h)
    hip.device_sync()
    out = np.empty_like(x)
    hip.dtoh(out, bufs.b.ptr)
    assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))
    for h, s in zip(handles, streams):
        assert L.librettDestroy(h) == L.LIBRETT_SUCCESS
        hip.stream_destroy(s)
    bufs.free()


def test_execute_after_destroy(gpu_stream, bufs):
    r, h = L.librettPlan(2, [8, 8], [1, 0], 4, gpu_stream)
    assert r == L.LIBRETT_SUCCESS
    assert L.librettDestroy(h) == L.LIBRETT_SUCCESS
    assert L.librettExecute(h, bufs.a.ptr, bufs.b.ptr) == L.LIBRETT_INVALID_PLAN
    r, h = L.librettPlan(2, [8, 8], [1, 0], 4, gpu_stream)
    assert L.librettExecute(h, bufs.a.ptr, bufs.a.ptr) == L.LIBRETT_INVALID_PARAMETER
    L.librettDestroy(h)


NAMED_LARGE = [
    ([8192, 8192], [1, 0], 4),
    ([1024, 256, 256], [0, 2, 1], 4),
    ([2, 8, 16, 32, 64, 128], [5, 3, 1, 0, 2, 4], 4),
    ([4096, 8192], [1, 0], 8),
    ([512, 256, 256], [0, 2, 1], 8),
    ([4, 4, 8, 32, 32, 256], [2, 5, 0, 4, 1, 3], 8),
    ([5, 97, 211, 331], [3, 1, 0, 2], 8),
]




@pytest.mark.parametrize("dim,perm,elem", NAMED_LARGE)
def test_named_large_cases(gpu_stream, big_bufs, dim, perm, elem):
    say("case: large", dim, perm, elem)
    info = check(dim, perm, elem, gpu_stream, big_bufs, seed=99)
    say("   ->", info["desc"])


def test_cpp_api_program():
    exe = os.path.join(ROOT, "librett_amd", "librett_apitest")
    assert os.path.exists(exe), "build librett_apitest first"
    say("case: librett_apitest (mangled C++ API)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    say(r.stdout)
    assert r.returncode == 0 and "apitest OK" in r.stdout, r.stdout + r.stderr
