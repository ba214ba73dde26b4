"""Device allocator hook on the GPU: with a caller's allocator installed,
the batched engine (and its autotuner), the serving engine and the checker
take every device buffer from it and give each one back; results stay
bit-exact."""
import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import data, hip, workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class Counting:
    def __init__(self):
        self.live = {}
        self.allocs = 0
        self.frees = 0

    def alloc(self, n):
        p = hip.malloc(n)
        self.live[p] = n
        self.allocs += 1
        return p

    def free(self, p):
        assert p in self.live
        del self.live[p]
        self.frees += 1
        hip.free(p)


def test_engines_use_the_callers_allocator(gpu_stream):
    say("case: caller's allocator for batch, autotuner, server, checker")
    pool = Counting()
    assert L.set_allocator(pool.alloc, pool.free) == L.LIBRETT_SUCCESS
    try:
        rng = np.random.default_rng(3)
        probs = [W.random_problem(rng, 3, 5, 1000, 50000, [8, 16]) for _ in range(40)]
        ins = [data.uniform(int(np.prod(d)), e, i) for i, (d, p, e) in enumerate(probs)]
        a = [hip.DeviceBuffer(x.nbytes) for x in ins]
        b = [hip.DeviceBuffer(x.nbytes) for x in ins]
        for buf, x in zip(a, ins):
            hip.htod(buf.ptr, x)
        bt = L.Batch(probs, gpu_stream, tune=2)
        n_batch = pool.allocs
        assert n_batch >= 2 and L.allocator_live() == 2  # scratch came and went; descriptors remain
        bt.set_buffers([x.ptr for x in a], [y.ptr for y in b])
        assert bt.execute(use_graph=True) == L.LIBRETT_SUCCESS
        hip.stream_sync(gpu_stream)
        for (d, p, e), x, y in zip(probs, ins, b):
            out = np.empty_like(x)
            hip.dtoh(out, y.ptr)
            assert np.array_equal(out.view(np.uint8), oracle.transpose(d, p, e, x).view(np.uint8))
        bt.destroy()
        assert L.allocator_live() == 0
        srv = L.Server(64 << 20, slots=2, max_batch=8)
        assert L.allocator_live() >= 1
        d, p, e = [64, 48, 32], [2, 0, 1], 8
        x = data.uniform(int(np.prod(d)), e, 9)
        pin, pout = srv.alloc(x.nbytes), srv.alloc(x.nbytes)
        hip.htod(pin, x)
        srv.wait(srv.submit(d, p, e, pin, pout))
        L.check_fill(pin, x.nbytes, gpu_stream)
        hip.stream_sync(gpu_stream)
        srv.wait(srv.submit(d, p, e, pin, pout))
        assert L.check_transpose(pout, d, p, e, gpu_stream) == (0, -1)
        srv.free(pin)
        srv.free(pout)
        srv.destroy()
        assert L.allocator_live() == 0
        assert pool.allocs == pool.frees and not pool.live
        for buf in a + b:
            buf.free()
    finally:
        assert L.set_allocator() == L.LIBRETT_SUCCESS
