"""Serving engine on the GPU: results of requests run alone equal results of
the same requests run in a concurrent mix (and the oracle), the arena never
exceeds its budget and returns every admitted byte, and admission waits
when a tiny arena is full."""
import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import data, hip, workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def requests(n, seed):
    rng = np.random.default_rng(seed)
    return [W.random_problem(rng, 2, 6, 1000, 400000, [4, 8, 16]) for _ in range(n)]


def test_alone_equals_mix():
    reqs = requests(120, 5)
    say("case: serving alone vs mix,", len(reqs), "requests")
    srv = L.Server(2 << 30, slots=4, max_batch=32)
    ins, outs_alone, outs_mix, refs = [], [], [], []
    for i, (d, p, e) in enumerate(reqs):
        x = data.uniform(int(np.prod(d)), e, 1000 + i)
        pin = srv.alloc(x.nbytes)
        hip.htod(pin, x)
        ins.append(pin)
        outs_alone.append(srv.alloc(x.nbytes))
        outs_mix.append(srv.alloc(x.nbytes))
        refs.append(oracle.transpose(d, p, e, x))
    # Alone: one request in the engine at a time.
    for i, (d, p, e) in enumerate(reqs):
        t = srv.submit(d, p, e, ins[i], outs_alone[i])
        srv.wait(t)
    # Mix: all at once, overlapping micro-batches.
    tickets = [srv.submit(d, p, e, ins[i], outs_mix[i]) for i, (d, p, e) in enumerate(reqs)]
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    say("  stats", {k: st[k] for k in ("completed", "batches", "mean_batch", "p50_us", "p99_us")})
    assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
    assert st["batches"] < 2 * len(reqs)  # the mix really was batched
    for i, ref in enumerate(refs):
        a = np.empty_like(ref)
        b = np.empty_like(ref)
        hip.dtoh(a, outs_alone[i])
        hip.dtoh(b, outs_mix[i])
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), reqs[i]
        assert np.array_equal(a.view(np.uint8), ref.view(np.uint8)), reqs[i]
    for p in ins + outs_alone + outs_mix:
        assert srv.free(p) == L.LIBRETT_SUCCESS
    assert srv.stats()["arena_in_use_bytes"] == 0
    srv.destroy()


def test_client_threads_share_the_engine():
    # Four host threads allocate from the arena, submit, wait and free at
    # the same time (ctypes releases the GIL in the library's calls); every
    # result equals the oracle and the arena ends empty.
    import threading
    say("case: serving from 4 client threads, 40 requests each")
    srv = L.Server(1 << 30, slots=4, max_batch=16)
    errors = []

    def client(k):
        try:
            for i, (d, p, e) in enumerate(requests(40, 100 + k)):
                x = data.uniform(int(np.prod(d)), e, 10000 * k + i)
                pin, pout = srv.alloc(x.nbytes), srv.alloc(x.nbytes)
                assert pin is not None and pout is not None
                hip.htod(pin, x)
                srv.wait(srv.submit(d, p, e, pin, pout))
                y = np.empty_like(x)
                hip.dtoh(y, pout)
                if not np.array_equal(y.view(np.uint8), oracle.transpose(d, p, e, x).view(np.uint8)):
                    errors.append((k, i, d, p, e))
                assert srv.free(pin) == L.LIBRETT_SUCCESS and srv.free(pout) == L.LIBRETT_SUCCESS
        except Exception as ex:  # reported below, from the main thread
            errors.append((k, repr(ex)))

    ts = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    st = srv.stats()
    say("  stats", {k: st[k] for k in ("completed", "batches", "mean_batch", "failed")})
    assert not errors, errors[:3]
    assert st["completed"] == 160 and st["failed"] == 0
    assert st["arena_in_use_bytes"] == 0
    srv.destroy()


def test_small_arena_admission_and_accounting():
    say("case: serving with a 64 MiB arena")
    budget = 64 << 20
    srv = L.Server(budget, slots=2, max_batch=16)
    pool = hip.DeviceBuffer(16 << 20)
    hip.memset(pool.ptr, 3, 16 << 20)
    st = srv.run_poisson(pool.ptr, 16 << 20, qps=200000, requests=3000, seed=9, catalog=64, min_bytes=4096,
                         max_bytes=8 << 20)
    say("  stats", st)
    assert st["completed"] == 3000 and st["failed"] == 0
    assert st["arena_peak_bytes"] <= budget
    assert st["arena_in_use_bytes"] == 0
    assert st["admission_stalls"] > 0
    assert srv.alloc(budget + 4096) is None  # larger than the arena: refused
    srv.destroy()
    pool.free()


def test_full_size_requests_checked_on_device(gpu_stream):
    # Config-4 sizes: 4 KiB - 64 MiB log-uniform, rank 2-6, f32/f64/c128,
    # inputs and outputs from the arena; every output checked on the device.
    rng = np.random.default_rng(44)
    srv = L.Server(16 << 30, slots=8, max_batch=64)
    reqs = []
    for _ in range(240):
        e = int(rng.choice([4, 8, 16]))
        nb = float(np.exp(rng.uniform(np.log(4096), np.log(64 << 20))))
        r = int(rng.integers(2, 7))
        reqs.append(W.random_problem(rng, r, r, max(2.0, nb / e), max(2.0, nb / e), [e]))
    reqs.append(([2048, 2048, 4], [1, 0, 2], 4))  # exactly 64 MiB
    say("case: %d full-size serving requests, largest %.1f MiB" % (
        len(reqs), max(np.prod(d) * e for d, _, e in reqs) / 2 ** 20))
    for wave in range(0, len(reqs), 80):
        part = reqs[wave:wave + 80]
        bufs = []
        for d, p, e in part:
            n = int(np.prod(d)) * e
            pin, pout = srv.alloc(n), srv.alloc(n)
            assert pin is not None and pout is not None
            L.check_fill(pin, n, gpu_stream)
            bufs.append((pin, pout))
        hip.stream_sync(gpu_stream)
        tickets = [srv.submit(d, p, e, pin, pout) for (d, p, e), (pin, pout) in zip(part, bufs)]
        for t in tickets:
            srv.wait(t)
        for (d, p, e), (pin, pout) in zip(part, bufs):
            assert L.check_transpose(pout, d, p, e, gpu_stream) == (0, -1), (d, p, e)
            assert srv.free(pin) == L.LIBRETT_SUCCESS and srv.free(pout) == L.LIBRETT_SUCCESS
    st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == len(reqs)
    assert st["arena_in_use_bytes"] == 0
    srv.destroy()


def test_poisson_fresh_shapes_inputs_in_arena():
    # Every request a new shape (planned on arrival by one of four client
    # threads); inputs and outputs both reserved in the arena.
    say("case: Poisson, fresh shapes, inputs in the arena, 4 clients")
    srv = L.Server(4 << 30, slots=8, max_batch=64)
    pool = hip.DeviceBuffer(64 << 20)
    st = srv.run_poisson(pool.ptr, 64 << 20, qps=20000, requests=2000, seed=12, min_bytes=4096,
                         max_bytes=64 << 20, fresh=True, inputs_in_arena=True, clients=4)
    say("  stats", {k: st[k] for k in ("completed", "plan_misses", "plan_seconds", "p50_us", "p99_us", "gbs",
                                        "arena_peak_bytes")})
    assert st["completed"] == 2000 and st["failed"] == 0
    assert st["plan_misses"] >= 1900  # random shapes rarely repeat
    assert st["plan_seconds"] > 0
    assert st["arena_in_use_bytes"] == 0
    assert st["arena_peak_bytes"] <= 4 << 30
    srv.destroy()
    pool.free()
