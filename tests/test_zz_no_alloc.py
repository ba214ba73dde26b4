"""The allocation-free checker entry point (librett_c_check_transpose_noalloc)
gives the same answer as librett_c_check_transpose on the same outputs, on
two calls in a row (the second smaller, so stale scratch from the first is
present), with correct outputs and with planted wrong elements, and takes
nothing from the device allocator."""
import ctypes

import numpy as np
import pytest

import librett_amd as L
from librett_amd import hip

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def noalloc(optr, dim, perm, elem, stream, res):
    assert L.check_transpose_noalloc(optr, dim, perm, elem, stream, res.ptr) == L.LIBRETT_SUCCESS
    hip.stream_sync(stream)
    out = np.empty(2, dtype=np.uint64)
    hip.dtoh(out, res.ptr)
    n, first = int(out[0]), int(out[1])
    return n, (-1 if first == 0xFFFFFFFFFFFFFFFF else first)


class Counting:
    def __init__(self):
        self.allocs = 0

    def alloc(self, n):
        self.allocs += 1
        return hip.malloc(n)

    def free(self, p):
        hip.free(p)


def test_noalloc_checker_matches_the_allocating_one(gpu_stream):
    cases = [([1024, 256, 96], [2, 0, 1], 8), ([45, 67, 13], [1, 2, 0], 16)]  # second call smaller
    a = hip.DeviceBuffer(1024 * 256 * 96 * 8)
    b = hip.DeviceBuffer(1024 * 256 * 96 * 8)
    res = hip.DeviceBuffer(256)
    try:
        for dim, perm, elem in cases:
            vol = int(np.prod(dim))
            say("case: noalloc checker", dim, perm, elem)
            L.check_fill(a.ptr, vol * elem, gpu_stream)
            r, h = L.librettPlan(len(dim), dim, perm, elem, gpu_stream)
            assert r == L.LIBRETT_SUCCESS
            assert L.librettExecute(h, a.ptr, b.ptr) == L.LIBRETT_SUCCESS
            hip.stream_sync(gpu_stream)
            L.librettDestroy(h)
            # Correct output.
            ref = L.check_transpose(b.ptr, dim, perm, elem, gpu_stream)
            assert ref == (0, -1)
            pool = Counting()
            assert L.set_allocator(pool.alloc, pool.free) == L.LIBRETT_SUCCESS
            try:
                got = noalloc(b.ptr, dim, perm, elem, gpu_stream, res)
            finally:
                assert L.set_allocator() == L.LIBRETT_SUCCESS
            assert pool.allocs == 0
            assert got == ref, (got, ref)
            # Planted wrong elements: a run of bytes in the middle and the last element.
            nbytes = vol * elem
            hip.memset(b.ptr + (nbytes // 3) // elem * elem, 0x5A, 3 * elem)
            hip.memset(b.ptr + nbytes - elem, 0x11, elem)
            ref = L.check_transpose(b.ptr, dim, perm, elem, gpu_stream)
            assert ref[0] == 4
            got = noalloc(b.ptr, dim, perm, elem, gpu_stream, res)
            assert got == ref, (got, ref)
        # Refused arguments return an error and launch nothing.
        assert L.check_transpose_noalloc(b.ptr, [4, 4], [0, 0], 8, gpu_stream, res.ptr) == \
            L.LIBRETT_INVALID_PARAMETER
        assert L.check_transpose_noalloc(b.ptr, [4, 4], [1, 0], 8, gpu_stream, 0) == L.LIBRETT_INVALID_PARAMETER
    finally:
        a.free()
        b.free()
        res.free()
