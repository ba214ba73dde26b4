"""Access counters on the CPU.

Three checks pin the counters for every candidate plan of small problems:

* the counters' lane enumeration equals, access for access, what the
  transpose kernel's own code does when replayed on the host (every load
  and store of every thread, in order, with its byte offset and size);
* the library's counts equal counts taken by definition from that access
  list (``oracle.counters``: ``numpy.unique`` over explicit byte positions);
* each byte of the tensor is loaded once and stored once (bytes counts),
  and every work item is counted.

Where the planner's model walks every work item (Tiled and TiledCopy with
at most 24 items), its line counts must equal the exact ones.
"""
import numpy as np
import pytest

import librett_amd as L
from oracle import counters as OC

PROBLEMS = [
    ((64, 48, 32), (2, 0, 1), 8),     # reference bench config 1 (plain)
    ((33, 70), (1, 0), 4),            # partial tiles both ways
    ((130, 5, 17), (0, 2, 1), 4),     # TiledCopy, odd words
    ((7, 9, 11, 5), (3, 1, 0, 2), 8),
    ((12, 20, 6), (1, 2, 0), 16),
    ((128, 96), (1, 0), 4),           # vector accesses, full tiles
    ((5, 97, 21), (2, 1, 0), 8),
    ((2, 3, 4, 5, 6), (4, 2, 0, 3, 1), 4),
    ((40, 3, 50), (0, 2, 1), 16),
    ((256, 8, 9), (0, 2, 1), 8),
    ((34, 18, 3), (1, 0, 2), 4),      # 8-byte f32 lanes
    ((20, 31, 17, 5), (3, 1, 0, 2), 4),  # Packed with 12 and 16 elements per thread (pipelined rounds)
]


def kernel_candidates(dim, perm, elem):
    return [(i, c) for i, c in enumerate(L.host_candidates(dim, perm, elem)) if c["method"] != 1]


def sorted_accesses(rec):
    k = np.lexsort((rec[:, 2], rec[:, 3], rec[:, 1], rec[:, 0]))
    return rec[k][:, [0, 1, 3, 4, 5]]


@pytest.mark.parametrize("dim,perm,elem", PROBLEMS)
def test_enumeration_is_the_kernels_accesses(dim, perm, elem):
    cands = kernel_candidates(dim, perm, elem)
    assert cands
    for i, c in cands:
        mine = L.host_access_trace(dim, perm, elem, i, 0)
        kern = L.host_access_trace(dim, perm, elem, i, 1)
        assert mine.shape == kern.shape, c["desc"]
        assert np.array_equal(sorted_accesses(mine), sorted_accesses(kern)), c["desc"]


@pytest.mark.parametrize("dim,perm,elem", PROBLEMS)
def test_counts_match_definition(dim, perm, elem):
    vol = int(np.prod(dim))
    for i, c in kernel_candidates(dim, perm, elem):
        got = L.host_candidate_counters(dim, perm, elem, i)
        ref = OC.counts(L.host_access_trace(dim, perm, elem, i, 1))
        for k, v in ref.items():
            assert got[k] == v, (c["desc"], k, got[k], v)
        assert got["ld_bytes"] == got["st_bytes"] == vol * elem, c["desc"]
        assert got["items"] == int(c["feat"][2]), c["desc"]
        assert got["ld_tran"] >= got["ld_req"] and got["st_tran"] >= got["st_req"]
        assert got["write_lines"] >= -(-vol * elem // 128) and got["partial_lines"] <= got["write_lines"]


def test_model_lines_exact_when_every_item_is_sampled():
    seen = 0
    for dim, perm, elem in PROBLEMS:
        for i, c in kernel_candidates(dim, perm, elem):
            if c["method"] not in (4, 5) or c["feat"][2] > 24:
                continue
            got = L.host_candidate_counters(dim, perm, elem, i)
            assert got["read_lines"] == c["read_lines"], c["desc"]
            assert got["write_lines"] == c["write_lines"], c["desc"]
            assert got["partial_lines"] == c["partial_lines"], c["desc"]
            seen += 1
    assert seen >= 5


def test_closed_form_full_tile():
    # 128 x 96 f32, 64 x 64 tiles with 16-byte lanes: each load instruction
    # of a wave covers 4 input rows of 64 floats (2 segments each), each
    # store instruction 4 output columns of 64 floats.
    dim, perm, elem = (128, 96), (1, 0), 4
    for i, c in kernel_candidates(dim, perm, elem):
        if c["method"] == 4 and (c["tile_x"], c["tile_y"], c["vec_x"], c["vec_y"]) == (64, 64, 4, 4):
            got = L.host_candidate_counters(dim, perm, elem, i)
            tiles = 2 * 2  # 128 / 64 along x, 96 / 64 -> 2 along y (one partial)
            assert got["items"] == tiles
            assert got["ld_bytes"] == 128 * 96 * 4
            # Input rows: 96 rows of 512 B = 4 lines each; output: 128 rows of 384 B = 3 lines each.
            assert got["read_lines"] == 96 * 4
            assert got["write_lines"] == 128 * 3
            assert got["partial_lines"] == 0
            return
    pytest.fail("64x64 vec 4 tile not among the candidates")


def test_planted_mistake_is_caught():
    dim, perm, elem = (33, 70), (1, 0), 4
    i, _ = kernel_candidates(dim, perm, elem)[0]
    rec = L.host_access_trace(dim, perm, elem, i, 1)
    good = OC.counts(rec)
    assert good == {k: v for k, v in L.host_candidate_counters(dim, perm, elem, i).items() if k in good}
    bad = rec.copy()
    j = np.flatnonzero(bad[:, 1] == 1)[0]
    bad[j, 4] += 128 * 7  # one store lands 7 lines away
    wrong = OC.counts(bad)
    assert wrong != good


def test_errors():
    with pytest.raises(RuntimeError):
        L.host_candidate_counters((8, 8), (0, 1), 4, 0)  # identity: Trivial copy, no kernel
    with pytest.raises(RuntimeError):
        L.host_candidate_counters((8, 8), (1, 0), 4, 999)
    with pytest.raises(RuntimeError):
        L.host_access_trace((8, 8), (1, 0), 4, 0, 2)


def test_by_definition_counts_hand_worked():
    # Accesses (item, store, instruction, thread, byte offset, bytes) whose
    # counts are worked out by hand, including lines written only in part
    # and a line completed by two instructions of one item.
    rec = []
    for t in range(64):  # item 0: one load, bytes 0..255 -> 2 segments, 2 lines
        rec.append((0, 0, 0, t, 4 * t, 4))
    for t in range(64):  # item 0: one store, bytes 64..319 -> lines 0 (64 B), 1 (full), 2 (64 B)
        rec.append((0, 1, 0, t, 64 + 4 * t, 4))
    for t in range(16):  # item 1, wave 0: 16-byte lanes, bytes 1000..1255 -> lines 7 (24 B), 8, 9 (104 B)
        rec.append((1, 1, 0, t, 1000 + 16 * t, 16))
    for t in range(64, 128):  # item 1, wave 1: bytes 2048..2303 -> lines 16, 17, both full
        rec.append((1, 1, 0, t, 2048 + 4 * (t - 64), 4))
    for t in range(16):  # item 2: two instructions fill line 32 together
        rec.append((2, 1, 0, t, 4096 + 4 * t, 4))
        rec.append((2, 1, 1, t, 4160 + 4 * t, 4))
    got = OC.counts(np.array(rec))
    assert got == {"ld_req": 1, "ld_tran": 2, "ld_bytes": 256, "read_lines": 2,
                   "st_req": 5, "st_tran": 3 + 3 + 2 + 1 + 1, "st_bytes": 256 + 256 + 256 + 128,
                   "write_lines": 3 + 3 + 2 + 1, "partial_lines": 2 + 2 + 0 + 0, "items": 3}
