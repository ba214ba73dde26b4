"""Access counters on the GPU: the device counting kernel against the host
run of the same code and against counts taken by definition from the
kernel's own accesses (small problems), and the invariants at full size
(every byte loaded once and stored once, every work item counted) for the
seven ~256 MiB bench problems and seeded random ones, next to the planner's
sampled estimates."""
import time

import numpy as np
import pytest

import librett_amd as L
from librett_amd import workloads as W
from oracle import counters as OC

pytestmark = pytest.mark.gpu

SMALL = [
    ((64, 48, 32), (2, 0, 1), 8),
    ((33, 70), (1, 0), 4),
    ((130, 5, 17), (0, 2, 1), 4),
    ((7, 9, 11, 5), (3, 1, 0, 2), 8),
    ((12, 20, 6), (1, 2, 0), 16),
    ((128, 96), (1, 0), 4),
    ((2, 3, 4, 5, 6), (4, 2, 0, 3, 1), 4),
]

LARGE = [
    ((8192, 8192), (1, 0), 4),
    ((1024, 256, 256), (0, 2, 1), 4),
    ((2, 8, 16, 32, 64, 128), (5, 3, 1, 0, 2, 4), 4),
    ((4096, 8192), (1, 0), 8),
    ((512, 256, 256), (0, 2, 1), 8),
    ((4, 4, 8, 32, 32, 256), (2, 5, 0, 4, 1, 3), 8),
    ((5, 97, 211, 331), (3, 1, 0, 2), 8),
]


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("dim,perm,elem", SMALL)
def test_device_counts_equal_host_and_definition(gpu_stream, dim, perm, elem):
    say("case: counters", dim, perm, elem)
    host = {c["desc"]: i for i, c in enumerate(L.host_candidates(dim, perm, elem))}
    matched = 0
    for i, c in enumerate(L.candidates(dim, perm, elem)):
        if c["method"] == 1 or c["desc"] not in host:
            continue
        j = host[c["desc"]]
        dev = L.candidate_counters(dim, perm, elem, i, gpu_stream)
        ref = L.host_candidate_counters(dim, perm, elem, j)
        assert dev == ref, c["desc"]
        byd = OC.counts(L.host_access_trace(dim, perm, elem, j, 1))
        for k, v in byd.items():
            assert dev[k] == v, (c["desc"], k)
        matched += 1
    assert matched >= 1


def check_full_size(dim, perm, elem, stream):
    vol = int(np.prod(dim))
    r, h = L.librettPlan(len(dim), list(dim), list(perm), elem, stream)
    assert r == 0
    try:
        info = L.plan_info(h)
        if info["method"] == 1:
            return None
        t0 = time.perf_counter()
        got = L.plan_counters(h)
        dt = time.perf_counter() - t0
    finally:
        L.librettDestroy(h)
    assert got["ld_bytes"] == got["st_bytes"] == vol * elem
    assert got["items"] == int(info["feat"][2])
    assert got["ld_tran"] >= got["ld_req"] > 0 and got["st_tran"] >= got["st_req"] > 0
    assert -(-vol * elem // 128) <= got["write_lines"]
    ratios = (info["read_lines"] / got["read_lines"], info["write_lines"] / got["write_lines"],
              info["feat"][0] / got["ld_tran"], info["feat"][1] / got["st_tran"])
    say("  %-30s %-28s %2d B %-12s exact lines r/w %d/%d tran %d/%d; model/exact lines %.3f/%.3f tran %.3f/%.3f "
        "(%.1f ms)" % (dim, perm, elem, info["method_name"], got["read_lines"], got["write_lines"], got["ld_tran"],
                       got["st_tran"], *ratios, dt * 1e3))
    if info["method"] in (4, 5):
        # Tiled and TiledCopy count lines the same way as the model, which
        # samples 24 tiles: the estimate is near, not exact, when the outer
        # dims make tiles unlike (up to about a third off on the seeded problems).
        assert 0.5 < ratios[0] < 2.0 and 0.5 < ratios[1] < 2.0
    return ratios


def test_full_size_bench_problems(gpu_stream):
    say("case: counters, bench problems")
    for dim, perm, elem in LARGE:
        check_full_size(dim, perm, elem, gpu_stream)


def test_full_size_seeded_problems(gpu_stream):
    say("case: counters, seeded random ~256 MiB problems")
    for seed in range(4):
        for elem in (4, 8):
            for cls in ("keep", "move", "small"):
                dim, perm = W.large_case(20261202 + seed, elem, cls)
                check_full_size(dim, perm, elem, gpu_stream)


def test_plan_with_exact_counts_is_correct(gpu_stream):
    # LIBRETT_EXACT_COUNTS=1 makes librettPlan count every sampled Tiled /
    # TiledCopy candidate on the device before choosing; the plan it picks
    # must still transpose bit-exactly. Run in a fresh process (the variable
    # is read once).
    import os
    import subprocess
    import sys
    say("case: librettPlan with LIBRETT_EXACT_COUNTS=1, (300, 200, 9)/(1, 2, 0) f32 and (1300, 700)/(1, 0) f64")
    code = r'''
import numpy as np, librett_amd as L, oracle
from librett_amd import data, hip
s = hip.stream_create()
for dim, perm, elem in (([300, 200, 9], [1, 2, 0], 4), ([1300, 700], [1, 0], 8)):
    n = int(np.prod(dim))
    x = data.uniform(n, elem, 5)
    a, b = hip.DeviceBuffer(n * elem), hip.DeviceBuffer(n * elem)
    hip.htod(a.ptr, x)
    r, h = L.librettPlan(len(dim), dim, perm, elem, s)
    assert r == 0
    assert L.librettExecute(h, a.ptr, b.ptr) == 0
    hip.stream_sync(s)
    y = np.empty_like(x)
    hip.dtoh(y, b.ptr)
    assert np.array_equal(y.view(np.uint8), oracle.transpose(dim, perm, elem, x).view(np.uint8))
    print(L.plan_info(h)["desc"], flush=True)
    L.librettDestroy(h)
print("ok")
'''
    env = dict(os.environ, LIBRETT_EXACT_COUNTS="1")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", code], env=env, cwd=root, capture_output=True, text=True, timeout=120)
    say(out.stdout, out.stderr[-2000:])
    assert out.returncode == 0 and out.stdout.strip().endswith("ok")
