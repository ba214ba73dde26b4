"""Brute-force access counts from a list of accesses -- test infrastructure only.

Only ``tests/`` may import this module, and only as the checker of the
library's access counters (``librett_c_*_counters``).

Each access is (work item, 0 load / 1 store, instruction, thread, byte
offset, bytes). The definitions follow the reference's MemStat counters
(``src/GpuModelKernel.cpp:43-64``): requests are wave instructions
(``gld_req`` / ``gst_req``), transactions are the distinct segments one wave
instruction touches (``countGlTransactions``, ``:95-118``, there counted as
segment changes between neighbouring lanes), here with 128-byte segments and
64-lane waves; lines are 128-byte lines. The reference counts full and
partial cache lines per warp instruction (``countCacheLines``,
``:136-173``); the library counts them per work item, so this module does
too: the distinct lines an item reads or writes, and the written lines whose
bytes the item covers only in part.

Everything is computed by definition with ``numpy.unique`` over the explicit
byte positions, with no run merging and no ballots, so it is independent of
how the library walks lanes and runs. Sized for small problems.
"""
import numpy as np


def counts(records):
    """Dict of counts from an (n, 6) int64 array of accesses."""
    rec = np.asarray(records, dtype=np.int64).reshape(-1, 6)
    item, store, inst, thread, off, nb = rec.T
    out = {}
    for s, pre in ((0, "ld"), (1, "st")):
        m = store == s
        if not m.any():
            out.update({pre + "_req": 0, pre + "_tran": 0, pre + "_bytes": 0})
            out["read_lines" if s == 0 else "write_lines"] = 0
            if s == 1:
                out["partial_lines"] = 0
            continue
        # One wave instruction = (item, instruction, thread // 64); keys are
        # packed into one int64 (items < 2^31, instructions < 2^8, waves
        # < 2^8, lines < 2^28).
        keys = (item[m] << 16) | (inst[m] << 8) | (thread[m] // 64)
        _, inv = np.unique(keys, return_inverse=True)
        inv = inv.ravel().astype(np.int64)
        out[pre + "_req"] = int(inv.max()) + 1
        # Every byte's segment, deduplicated per instruction.
        o, n = off[m], nb[m]
        rep = np.repeat(np.arange(o.size), n)
        pos = o[rep] + (np.arange(rep.size) - np.repeat(np.cumsum(n) - n, n))
        out[pre + "_tran"] = int(np.unique((inv[rep] << 32) | (pos >> 7)).size)
        out[pre + "_bytes"] = int(n.sum())
        # Lines per item; coverage counts bytes (each byte is accessed once
        # per item).
        lines, cover = np.unique((item[m][rep] << 32) | (pos >> 7), return_counts=True)
        if s == 0:
            out["read_lines"] = int(lines.size)
        else:
            out["write_lines"] = int(lines.size)
            out["partial_lines"] = int((cover < 128).sum())
    out["items"] = int(np.unique(item).size)
    return out
