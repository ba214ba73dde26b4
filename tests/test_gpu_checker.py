"""Index-pattern checker on the GPU: agreement with the host replay and the
oracle, planted errors, and the seeded random ~256 MiB transposes (keep /
move / small leading dim, f32 and f64, 8 seeds each) checked on the device
for the model's and the measured plans."""
import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import hip
from librett_amd import workloads as W

pytestmark = pytest.mark.gpu

CAP = 300 << 20


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def bufs():
    a, b = hip.DeviceBuffer(CAP), hip.DeviceBuffer(CAP)
    yield a, b
    a.free()
    b.free()


def test_fill_and_check_small(gpu_stream, bufs):
    a, b = bufs
    say("case: checker fill + planted errors, (7,5,11,3)/(2,0,3,1)")
    d, p = [7, 5, 11, 3], [2, 0, 3, 1]
    for elem in (4, 8, 16):
        n = int(np.prod(d)) * elem
        L.check_fill(a.ptr, n, gpu_stream)
        hip.stream_sync(gpu_stream)
        x = np.zeros(n // 4, dtype=np.uint32)
        hip.dtoh(x, a.ptr)
        ref = np.zeros_like(x)
        L.host_check_fill(ref)
        assert np.array_equal(x, ref)
        y = oracle.transpose(d, p, elem, x.view(oracle.ELEM_DTYPES[elem])).view(np.uint32).copy()
        hip.htod(b.ptr, y)
        assert L.check_transpose(b.ptr, d, p, elem, gpu_stream) == (0, -1)
        y[[5, 17, 40]] += 1
        hip.htod(b.ptr, y)
        words = elem // 4
        bad = sorted(set(w // words for w in (5, 17, 40)))
        got = L.check_transpose(b.ptr, d, p, elem, gpu_stream)
        assert got == (len(bad), bad[0]) == L.host_check_transpose(y, d, p, elem)


CASES = [(seed, elem, cls) for elem in (4, 8) for cls in ("keep", "move", "small") for seed in range(8)]


@pytest.mark.parametrize("seed,elem,cls", CASES)
def test_seeded_large_cases(gpu_stream, bufs, seed, elem, cls):
    a, b = bufs
    d, p = W.large_case(20261202 + seed, elem, cls)
    n = int(np.prod(d)) * elem
    assert n <= CAP
    say("case: large %s f%d seed %d dims %s perm %s" % (cls, 8 * elem, seed, d, p))
    L.check_fill(a.ptr, n, gpu_stream)
    modes = ("plan", "measure") if seed < 2 else ("plan",)
    for mode in modes:
        hip.memset(b.ptr, 0xEE, n)
        if mode == "plan":
            r, h = L.librettPlan(len(d), d, p, elem, gpu_stream)
        else:
            r, h = L.librettPlanMeasure(len(d), d, p, elem, gpu_stream, a.ptr, b.ptr)
        assert r == L.LIBRETT_SUCCESS
        hip.memset(b.ptr, 0xEE, n)
        assert L.librettExecute(h, a.ptr, b.ptr) == L.LIBRETT_SUCCESS
        got = L.check_transpose(b.ptr, d, p, elem, gpu_stream)
        say("   %s -> %s: %s" % (mode, L.plan_info(h)["desc"], got))
        assert got == (0, -1)
        L.librettDestroy(h)
