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
