"""Writes ttc_cases.json: the (dims, permutation) pairs of the reference's
TTC benchmark (tests/librett_bench.cpp:576-708, bench6, taken from "TTC: A
Tensor Transposition Compiler for Multiple Architectures") and its bench7
rank-8 and rank-12 shapes (tests/librett_bench.cpp:713-757).

The pairs are read as data from the reference source text; nothing of the
reference is run. Usage: python tests/golden/make_ttc_cases.py [REF_DIR]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def int_vectors(block):
    return [[int(v) for v in m.split(",")] for m in re.findall(r"std::vector<int>\{([0-9,\s]+)\}", block)]


def main(ref=sys.argv[1] if len(sys.argv) > 1 else "/root/reference"):
    src = open(os.path.join(ref, "tests", "librett_bench.cpp")).read()
    b6 = src[src.index("bool bench6(gpuStream_t& gpuStream) {"):]
    b6 = b6[:b6.index("for (int i=0;i < dims.size();i++)")]
    dims_block = b6[b6.index("dims = {"):b6.index("permutations = {")]
    perm_block = b6[b6.index("permutations = {"):]
    dims, perms = int_vectors(dims_block), int_vectors(perm_block)
    assert len(dims) == len(perms) == 57, (len(dims), len(perms))
    for d, p in zip(dims, perms):
        assert sorted(p) == list(range(len(d))), (d, p)
    b7 = src[src.index("bool bench7(gpuStream_t& gpuStream) {"):]
    b7 = b7[:b7.index("return true;")]
    shapes7 = [[int(v) for v in m.split(",")] for m in re.findall(r"std::vector<int> dim = \{([0-9,\s]+)\};", b7)]
    assert [len(s) for s in shapes7] == [8, 12], shapes7
    out = {
        "source": "tests/librett_bench.cpp:576-708 (bench6, TTC suite, 8-byte elements) and :713-757 (bench7)",
        "ttc": [{"dim": d, "perm": p} for d, p in zip(dims, perms)],
        "bench7": shapes7,
    }
    with open(os.path.join(HERE, "ttc_cases.json"), "w") as fh:
        fh.write("{\"source\": %s,\n \"ttc\": [\n" % json.dumps(out["source"]))
        fh.write(",\n".join("  " + json.dumps(c) for c in out["ttc"]))
        fh.write("\n ],\n \"bench7\": %s}\n" % json.dumps(out["bench7"]))
    print("wrote %d TTC pairs and %d bench7 shapes" % (len(dims), len(shapes7)))


if __name__ == "__main__":
    main()
