"""Robustness of the library on the GPU: the copy-ceiling kernel's output,
the dynamic-LDS limit shared by plans of one kernel, and the API used from
many host threads at once.

Reference behaviour these cover: the bench copy kernels
(src/GpuMemcpy.cpp:141-209) and the plan map / handle counter races of
src/librett.cpp:128-135,375, which this library's API is built to avoid.
"""
import threading

import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import data, hip

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_device_copy_every_variant(gpu_stream):
    # Sizes that are not multiples of a block's 256*K words, explicit small
    # grids (grid-stride loop) and the default grid; bytes past the end stay.
    src = hip.DeviceBuffer(1 << 22)
    dst = hip.DeviceBuffer(1 << 22)
    rng = np.random.default_rng(11)
    x = rng.integers(0, 256, size=1 << 22, dtype=np.uint8)
    hip.htod(src.ptr, x)
    try:
        for k in (1, 2, 4, 8):
            for nt in (0, 1):
                for nbytes in (16, 4096 + 16, 256 * 16 * k * 3 + 48, (1 << 22) - 16 * 7):
                    for grid in (0, 1, 3, 37):
                        say("case: device_copy K=%d nt=%d bytes=%d grid=%d" % (k, nt, nbytes, grid))
                        hip.memset(dst.ptr, 0xA5, 1 << 22)
                        assert L.device_copy(dst.ptr, src.ptr, nbytes, gpu_stream, k, nt, grid) == L.LIBRETT_SUCCESS
                        hip.stream_sync(gpu_stream)
                        y = np.empty(1 << 22, dtype=np.uint8)
                        hip.dtoh(y, dst.ptr)
                        assert np.array_equal(y[:nbytes], x[:nbytes]), (k, nt, nbytes, grid)
                        assert np.all(y[nbytes:] == 0xA5), (k, nt, nbytes, grid)
        # Arguments the kernel cannot take are refused, not launched.
        assert L.device_copy(dst.ptr, src.ptr, 20, gpu_stream, 1, 0, 0) == L.LIBRETT_INVALID_PARAMETER
        assert L.device_copy(dst.ptr + 8, src.ptr, 32, gpu_stream, 1, 0, 0) == L.LIBRETT_INVALID_PARAMETER
        assert L.device_copy(dst.ptr, src.ptr, 32, gpu_stream, 3, 0, 0) == L.LIBRETT_INVALID_PARAMETER
    finally:
        src.free()
        dst.free()


def _run_and_check(h, dim, perm, elem, x, a, b, stream):
    hip.htod(a.ptr, x)
    hip.memset(b.ptr, 0xEE, x.nbytes)
    assert L.librettExecute(h, a.ptr, b.ptr) == L.LIBRETT_SUCCESS
    hip.stream_sync(stream)
    out = np.empty_like(x)
    hip.dtoh(out, b.ptr)
    ref = oracle.transpose(dim, perm, elem, x)
    assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))


def test_lds_limit_is_not_lowered_by_a_later_plan(gpu_stream):
    # Plan A stages more than 64 KiB of LDS in k_packed<8,16>; plan B uses the
    # same kernel with less. Creating B after A must not break A.
    dim, perm, elem = [512, 256, 256], [0, 2, 1], 8
    cands = L.candidates(dim, perm, elem)
    big = [i for i, c in enumerate(cands) if c["method"] in (L.PACKED, L.PACKED_SPLIT) and c["nreg"] == 16
           and c["lds_bytes"] > 128 * 1024]
    small = [i for i, c in enumerate(cands) if c["method"] in (L.PACKED, L.PACKED_SPLIT) and c["nreg"] == 16
             and 64 * 1024 < c["lds_bytes"] < 100 * 1024]
    assert big and small, [c["desc"] for c in cands]
    say("case: LDS limit", cands[big[0]]["desc"], "then", cands[small[0]]["desc"])
    vol = int(np.prod(dim))
    x = data.uniform(vol, elem, 5)
    a, b = hip.DeviceBuffer(x.nbytes), hip.DeviceBuffer(x.nbytes)
    try:
        ra, ha = L.plan_candidate(len(dim), dim, perm, elem, gpu_stream, big[0])
        assert ra == L.LIBRETT_SUCCESS
        rb, hb = L.plan_candidate(len(dim), dim, perm, elem, gpu_stream, small[0])
        assert rb == L.LIBRETT_SUCCESS
        _run_and_check(ha, dim, perm, elem, x, a, b, gpu_stream)
        _run_and_check(hb, dim, perm, elem, x, a, b, gpu_stream)
        # Planning again (the candidate list queries occupancy for every LDS
        # size) leaves A runnable too.
        assert L.candidates(dim, perm, elem)
        _run_and_check(ha, dim, perm, elem, x, a, b, gpu_stream)
        for h in (ha, hb):
            assert L.librettDestroy(h) == L.LIBRETT_SUCCESS
    finally:
        a.free()
        b.free()


def test_many_threads_plan_execute_destroy():
    # Eight host threads, each with its own stream, plan, execute, check and
    # destroy their own problems while sharing two plans made up front.
    nthreads, per_thread = 8, 12
    shared_dim, shared_perm, shared_elem = [48, 40, 36], [2, 0, 1], 8
    shared_stream = hip.stream_create()
    r, shared = L.librettPlan(3, shared_dim, shared_perm, shared_elem, shared_stream)
    assert r == L.LIBRETT_SUCCESS
    errors = []
    handles = []
    lock = threading.Lock()

    def worker(tid):
        try:
            rng = np.random.default_rng(100 + tid)
            stream = hip.stream_create()
            a, b = hip.DeviceBuffer(8 << 20), hip.DeviceBuffer(8 << 20)
            sx = data.uniform(int(np.prod(shared_dim)), shared_elem, 7)
            sref = oracle.transpose(shared_dim, shared_perm, shared_elem, sx)
            try:
                for k in range(per_thread):
                    rank = int(rng.integers(2, 6))
                    elem = int(rng.choice([4, 8, 16]))
                    dim = [int(v) for v in rng.integers(2, 24, size=rank)]
                    while int(np.prod(dim)) * elem > (8 << 20):
                        dim[int(np.argmax(dim))] //= 2
                    perm = [int(v) for v in rng.permutation(rank)]
                    x = data.uniform(int(np.prod(dim)), elem, tid * 1000 + k)
                    r, h = L.librettPlan(rank, dim, perm, elem, stream)
                    assert r == L.LIBRETT_SUCCESS, (dim, perm, elem, r)
                    with lock:
                        handles.append(h)
                    hip.htod(a.ptr, x)
                    assert L.librettExecute(h, a.ptr, b.ptr) == L.LIBRETT_SUCCESS
                    hip.stream_sync(stream)
                    out = np.empty_like(x)
                    hip.dtoh(out, b.ptr)
                    ref = oracle.transpose(dim, perm, elem, x)
                    assert np.array_equal(out.view(np.uint8), ref.view(np.uint8)), (dim, perm, elem)
                    assert L.librettDestroy(h) == L.LIBRETT_SUCCESS
                    assert L.librettExecute(h, a.ptr, b.ptr) == L.LIBRETT_INVALID_PLAN
                    # The shared plan runs on its own stream; each thread
                    # checks its own output buffer.
                    hip.htod(a.ptr, sx)
                    assert L.librettExecute(shared, a.ptr, b.ptr) == L.LIBRETT_SUCCESS
                    hip.stream_sync(shared_stream)
                    out = np.empty_like(sx)
                    hip.dtoh(out, b.ptr)
                    assert np.array_equal(out.view(np.uint8), sref.view(np.uint8))
            finally:
                a.free()
                b.free()
                hip.stream_destroy(stream)
        except BaseException as e:  # reported by the main thread
            errors.append((tid, repr(e)))

    say("case: %d threads x %d plans" % (nthreads, per_thread))
    ts = [threading.Thread(target=worker, args=(i,)) for i in range(nthreads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in ts), "a worker thread did not finish"
    assert not errors, errors
    # Every handle was distinct (one atomic counter).
    assert len(set(handles)) == len(handles) == nthreads * per_thread
    assert shared not in handles
    assert L.librettDestroy(shared) == L.LIBRETT_SUCCESS
    hip.stream_destroy(shared_stream)


def test_destroy_while_other_threads_execute(gpu_stream):
    # Executes racing a destroy either run (the plan stays alive for the
    # launch in flight) or report an invalid plan afterwards; nothing else.
    dim, perm, elem = [64, 48, 32], [2, 0, 1], 8
    vol = int(np.prod(dim))
    a, b = hip.DeviceBuffer(vol * elem), hip.DeviceBuffer(vol * elem)
    r, h = L.librettPlan(3, dim, perm, elem, gpu_stream)
    assert r == L.LIBRETT_SUCCESS
    codes = []
    start = threading.Event()

    def runner():
        start.wait()
        for _ in range(2000):
            codes.append(L.librettExecute(h, a.ptr, b.ptr))

    ts = [threading.Thread(target=runner) for _ in range(4)]
    for t in ts:
        t.start()
    start.set()
    assert L.librettDestroy(h) == L.LIBRETT_SUCCESS
    for t in ts:
        t.join(timeout=120)
    hip.stream_sync(gpu_stream)
    assert set(codes) <= {L.LIBRETT_SUCCESS, L.LIBRETT_INVALID_PLAN}
    assert codes.count(L.LIBRETT_INVALID_PLAN) > 0
    a.free()
    b.free()
