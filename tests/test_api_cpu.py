"""C-ABI library checks that need no GPU: it loads, exports every symbol the
headers declare (the extern "C" twins and the C++-mangled librett.h API), and
returns the reference's error codes for invalid input
(src/librett.cpp:85-108, :237, :354, :377-379)."""
import ctypes
import os
import re
import subprocess

import pytest

import librett_amd as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_c_symbols():
    text = open(os.path.join(ROOT, "include", "librett_c.h")).read()
    return sorted(set(re.findall(r"\b(librett_c_[a-z_]+)\s*\(", text)))


def test_library_loads_and_versions():
    assert "gfx950" in L.version()


def test_exports_every_c_symbol():
    lib = L.lib()
    names = header_c_symbols()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n


def test_exports_mangled_cpp_api():
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True, check=True)
    syms = out.stdout
    # Itanium mangling of the reference signatures (src/librett.h:57-121):
    # hipStream_t is ihipStream_t*, passed by reference.
    expected = [
        "_Z17librettInitializev",
        "_Z15librettFinalizev",
        "_Z11librettPlanPjiPiS0_mRP12ihipStream_t",
        "_Z18librettPlanMeasurePjiPiS0_mRP12ihipStream_tPvS4_",
        "_Z14librettDestroyj",
        "_Z14librettExecutejPvS_",
    ]
    for e in expected:
        assert e in syms, e


BAD_INPUTS = [
    # (rank, dim, perm, sizeofType)
    (2, [4, 4], [1, 0], 2),
    (2, [4, 4], [1, 0], 12),
    (0, [4], [0], 4),
    (2, [4, 0], [1, 0], 8),
    (2, [4, -3], [1, 0], 8),
    (2, [4, 4], [0, 0], 8),
    (2, [4, 4], [0, 2], 8),
    (2, [4, 4], [-1, 0], 8),
    (3, [2048, 2048, 1024], [2, 1, 0], 4),  # 2^32 elements: beyond 32-bit indexing
]


@pytest.mark.parametrize("rank,dim,perm,elem", BAD_INPUTS)
def test_invalid_parameters(rank, dim, perm, elem):
    r, _ = L.librettPlan(rank, dim, perm, elem, 0)
    assert r == L.LIBRETT_INVALID_PARAMETER
    r, _ = L.librettPlanMeasure(rank, dim, perm, elem, 0, 0x1000, 0x2000)
    assert r == L.LIBRETT_INVALID_PARAMETER
    if rank == len(dim):
        with pytest.raises(ValueError):
            L.host_candidates(dim, perm, elem)


def test_measure_rejects_in_place():
    r, _ = L.librettPlanMeasure(2, [4, 4], [1, 0], 8, 0, 0x1000, 0x1000)
    assert r == L.LIBRETT_INVALID_PARAMETER


def test_unknown_handle():
    assert L.librettExecute(987654, 0x1000, 0x2000) == L.LIBRETT_INVALID_PLAN
    assert L.librettDestroy(987654) == L.LIBRETT_INVALID_PLAN
    with pytest.raises(ValueError):
        L.plan_info(987654)


def test_host_replay_rejects_bad_index():
    import numpy as np
    x = np.zeros(16, dtype=np.float32)
    with pytest.raises(ValueError):
        L.host_replay([4, 4], [1, 0], 4, 999, x)


def test_allocator_hook_arguments():
    import ctypes
    lib = L.lib()
    # alloc without free (and the reverse) is refused; both NULL restores hipMalloc.
    fa = L.ALLOC_FN(lambda n, c: None)
    ff = L.FREE_FN(lambda p, c: None)
    assert lib.librett_c_set_allocator(fa, L.FREE_FN(), None) == L.LIBRETT_INVALID_PARAMETER
    assert lib.librett_c_set_allocator(L.ALLOC_FN(), ff, None) == L.LIBRETT_INVALID_PARAMETER
    assert L.set_allocator(lambda n: 0, lambda p: None) == L.LIBRETT_SUCCESS
    assert L.allocator_live() == 0
    assert L.set_allocator() == L.LIBRETT_SUCCESS
