"""Batched engine on the GPU: every problem of a batch bit-exact against the
oracle, for graph replays (several in a row, with other device work between
them), plain launches, autotuned plans, unaligned buffers (fallback
descriptors) and re-pointed buffers."""
import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import data, hip, workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_problems(n, seed, vol_hi=60000):
    rng = np.random.default_rng(seed)
    probs = [W.random_problem(rng, 3, 5, 1000, vol_hi, [8, 16]) for _ in range(n)]
    probs += [([3, 4, 5], [0, 1, 2], 8), ([64, 48, 32], [2, 0, 1], 8), ([2, 8, 16, 32], [3, 1, 0, 2], 4)]
    return probs


class Slab:
    """Inputs and outputs of a batch carved from two device allocations."""

    def __init__(self, problems, offset=0, seed=0):
        self.sizes = [int(np.prod(d)) * e for d, p, e in problems]
        self.offs = []
        o = 0
        for s in self.sizes:
            self.offs.append(o)
            o += (s + 256 + 255) // 256 * 256
        self.total = o + 512
        self.a = hip.DeviceBuffer(self.total)
        self.b = hip.DeviceBuffer(self.total)
        self.offset = offset
        self.inputs = [data.uniform(int(np.prod(d)), e, seed + i) for i, (d, p, e) in enumerate(problems)]
        for i, x in enumerate(self.inputs):
            hip.htod(self.a.ptr + self.offs[i] + offset, x)
        self.refs = [oracle.transpose(d, p, e, x) for (d, p, e), x in zip(problems, self.inputs)]

    def in_ptrs(self):
        return [self.a.ptr + o + self.offset for o in self.offs]

    def out_ptrs(self):
        return [self.b.ptr + o + self.offset for o in self.offs]

    def check(self):
        bad = []
        for i, ref in enumerate(self.refs):
            out = np.empty_like(ref)
            hip.dtoh(out, self.b.ptr + self.offs[i] + self.offset)
            if not np.array_equal(out.view(np.uint8), ref.view(np.uint8)):
                bad.append(i)
        return bad

    def free(self):
        self.a.free()
        self.b.free()


def test_batch_graph_replays(gpu_stream):
    probs = make_problems(300, 1)
    say("case: batch of", len(probs), "problems, graph replays")
    slab = Slab(probs)
    b = L.Batch(probs, gpu_stream, tune=0)
    b.set_buffers(slab.in_ptrs(), slab.out_ptrs())
    say("  info", b.info())
    for rep in range(3):
        hip.memset(slab.b.ptr, 0xEE, slab.total)          # eager work between replays
        L.device_copy(slab.b.ptr, slab.a.ptr, 1 << 20, gpu_stream, 1, 1, 0)
        hip.stream_sync(gpu_stream)
        hip.memset(slab.b.ptr, 0xEE, slab.total)
        assert b.execute(use_graph=True) == L.LIBRETT_SUCCESS
        hip.stream_sync(gpu_stream)
        bad = slab.check()
        assert not bad, ("replay", rep, [probs[i] for i in bad[:5]])
    b.destroy()
    slab.free()


def test_batch_plain_launch_and_repoint(gpu_stream):
    probs = make_problems(120, 2)
    say("case: batch plain launch + re-pointed buffers")
    s1, s2 = Slab(probs, seed=100), Slab(probs, seed=200)
    b = L.Batch(probs, gpu_stream, tune=0)
    for slab in (s1, s2, s1):
        hip.memset(slab.b.ptr, 0xEE, slab.total)
        b.set_buffers(slab.in_ptrs(), slab.out_ptrs())
        assert b.execute(use_graph=False) == L.LIBRETT_SUCCESS
        hip.stream_sync(gpu_stream)
        assert not slab.check()
        hip.memset(slab.b.ptr, 0xEE, slab.total)
        assert b.execute(use_graph=True) == L.LIBRETT_SUCCESS
        hip.stream_sync(gpu_stream)
        assert not slab.check()
    b.destroy()
    s1.free()
    s2.free()


@pytest.mark.parametrize("offset", [8, 4])
def test_batch_unaligned_buffers(gpu_stream, offset):
    probs = [p for p in make_problems(80, 3) if p[2] >= offset]
    say("case: batch unaligned by", offset)
    slab = Slab(probs, offset=offset)
    b = L.Batch(probs, gpu_stream, tune=0)
    b.set_buffers(slab.in_ptrs(), slab.out_ptrs())
    for _ in range(2):
        hip.memset(slab.b.ptr, 0xEE, slab.total)
        assert b.execute(use_graph=True) == L.LIBRETT_SUCCESS
        hip.stream_sync(gpu_stream)
        assert not slab.check()
    b.destroy()
    slab.free()


def test_batch_autotuned(gpu_stream):
    probs = make_problems(60, 4)
    say("case: autotuned batch")
    slab = Slab(probs)
    b = L.Batch(probs, gpu_stream, tune=3)
    info = b.info()
    say("  info", info)
    assert info["tuned"] >= 1
    b.set_buffers(slab.in_ptrs(), slab.out_ptrs())
    hip.memset(slab.b.ptr, 0xEE, slab.total)
    assert b.execute(use_graph=True) == L.LIBRETT_SUCCESS
    hip.stream_sync(gpu_stream)
    assert not slab.check()
    b.destroy()
    slab.free()


def test_batch_autotuned_problem_over_256mib(gpu_stream):
    # The autotuner times clones of each problem in scratch buffers sized for
    # the largest problem of the batch; a 269.5 MB problem must fit them.
    probs = [([5, 97, 211, 331], [3, 1, 0, 2], 8), ([64, 48, 32], [2, 0, 1], 8), ([33, 7, 130], [2, 0, 1], 16)]
    say("case: autotuned batch with a 269.5 MB problem")
    slab = Slab(probs, seed=7)
    b = L.Batch(probs, gpu_stream, tune=2)
    info = b.info()
    say("  info", info)
    assert info["tuned"] >= 1
    b.set_buffers(slab.in_ptrs(), slab.out_ptrs())
    hip.memset(slab.b.ptr, 0xEE, slab.total)
    assert b.execute(use_graph=True) == L.LIBRETT_SUCCESS
    hip.stream_sync(gpu_stream)
    assert not slab.check()
    b.destroy()
    slab.free()


def test_autotuner_not_slower_than_model(gpu_stream):
    # The autotuner times the model's best candidates (the model's own choice
    # among them) inside the persistent kernel, so its batch may not be
    # slower than the model's batch beyond timing noise (5%).
    probs = make_problems(400, 9, vol_hi=200000)
    say("case: autotuned vs model batch timing,", len(probs), "problems")
    slab = Slab(probs, seed=3)
    times = {}
    for tune in (0, 3):
        b = L.Batch(probs, gpu_stream, tune=tune)
        b.set_buffers(slab.in_ptrs(), slab.out_ptrs())
        e0, e1 = hip.Event(), hip.Event()
        ts = []
        for _ in range(3):
            b.execute(use_graph=True)
        for _ in range(7):
            e0.record(gpu_stream)
            for _ in range(5):
                b.execute(use_graph=True)
            e1.record(gpu_stream)
            e1.synchronize()
            ts.append(e1.elapsed_ms(e0) / 5)
        times[tune] = sorted(ts)[len(ts) // 2]
        hip.memset(slab.b.ptr, 0xEE, slab.total)
        assert b.execute(use_graph=True) == L.LIBRETT_SUCCESS
        hip.stream_sync(gpu_stream)
        assert not slab.check()
        b.destroy()
    say("  model %.3f ms, autotuned %.3f ms" % (times[0], times[3]))
    assert times[3] <= 1.05 * times[0]
    slab.free()


def test_config3_full_batch_checked_on_device(gpu_stream):
    # The batched workload at its stated size: 4096 problems of 1e3-1e6
    # elements (f64 / c128), autotuned, one graph launch; every output
    # checked on the device against the index pattern of its input.
    probs = W.batched_problems(4096)
    sizes = [int(np.prod(d)) * e for d, p, e in probs]
    offs, o = [], 0
    for s_ in sizes:
        offs.append(o)
        o += (s_ + 255) // 256 * 256
    say("case: 4096-problem batch, %.2f GB per side" % (o / 1e9))
    A, B = hip.DeviceBuffer(o), hip.DeviceBuffer(o)
    try:
        for off, s_ in zip(offs, sizes):
            L.check_fill(A.ptr + off, s_, gpu_stream)
        hip.memset(B.ptr, 0xEE, o)
        hip.stream_sync(gpu_stream)
        b = L.Batch(probs, gpu_stream, tune=3)
        say("  info", b.info())
        b.set_buffers([A.ptr + x for x in offs], [B.ptr + x for x in offs])
        for rep in range(2):
            assert b.execute(use_graph=True) == L.LIBRETT_SUCCESS
            hip.stream_sync(gpu_stream)
            bad = [i for i, ((d, p, e), off) in enumerate(zip(probs, offs))
                   if L.check_transpose(B.ptr + off, d, p, e, gpu_stream)[0]]
            assert not bad, (rep, [probs[i] for i in bad[:5]])
            hip.memset(B.ptr, 0xEE, o)
        b.destroy()
    finally:
        A.free()
        B.free()


def test_many_tiny_problems_between_large_ones(gpu_stream):
    # Thousands of problems of a few elements between a few large ones: a
    # workgroup that strides over units by the grid then skips many problems
    # per unit, which the kernel's wave-parallel search (64 problems per
    # step) must get right at every boundary, including runs longer than 64
    # problems and the last problem of each family.
    rng = np.random.default_rng(77)
    probs = []
    for i in range(3000):
        if i % 700 == 350:
            probs.append(W.random_problem(rng, 3, 4, 200000, 400000, [8]))
        else:
            r = int(rng.integers(2, 5))
            d = [int(x) for x in rng.integers(1, 4, size=r)]
            p = [int(x) for x in rng.permutation(r)]
            probs.append((d, p, int(rng.choice([4, 8, 16]))))
    say("case: 3000 problems, %d of them large" % sum(1 for d, p, e in probs if np.prod(d) > 1000))
    slab = Slab(probs, seed=5)
    try:
        b = L.Batch(probs, gpu_stream)
        b.set_buffers(slab.in_ptrs(), slab.out_ptrs())
        assert b.execute(use_graph=False) == L.LIBRETT_SUCCESS
        hip.stream_sync(gpu_stream)
        bad = slab.check()
        assert not bad, [probs[i] for i in bad[:5]]
        b.destroy()
    finally:
        slab.free()
