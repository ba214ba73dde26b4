"""The committed fixtures of the reference's benchmark inputs are well
formed, and the seeded test2-style generator behaves like the reference's
(tests/librett_test.cpp:194-272): volumes within the bound, dims >= 2."""
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from librett_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


def test_ttc_fixture_well_formed():
    f = json.load(open(os.path.join(HERE, "golden", "ttc_cases.json")))
    assert len(f["ttc"]) == 57
    for c in f["ttc"]:
        assert sorted(c["perm"]) == list(range(len(c["dim"])))
        assert all(d >= 1 for d in c["dim"])
        assert math.prod(c["dim"]) < 2 ** 31
    # The TTC suite's first and last pairs (tests/librett_bench.cpp:579,635,639,695).
    assert f["ttc"][0] == {"dim": [7248, 7248], "perm": [1, 0]}
    assert f["ttc"][-1] == {"dim": [32, 5, 15, 15, 15, 112], "perm": [5, 4, 3, 2, 1, 0]}
    assert f["bench7"] == [[5, 3, 2, 4, 35, 33, 37, 40], [2, 3, 4, 3, 2, 2, 3, 2, 20, 18, 22, 24]]


@pytest.mark.skipif(not os.path.isdir("/root/reference/tests"), reason="reference sources not present")
def test_ttc_fixture_matches_its_script(tmp_path):
    # The script reads the reference source text and must write what is committed.
    import shutil
    script = os.path.join(HERE, "golden", "make_ttc_cases.py")
    work = tmp_path / "golden"
    work.mkdir()
    shutil.copy(script, work / "make_ttc_cases.py")
    r = subprocess.run([sys.executable, str(work / "make_ttc_cases.py")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert (work / "ttc_cases.json").read_text() == open(os.path.join(HERE, "golden", "ttc_cases.json")).read()


def test_test2_dims_generator():
    rng = np.random.default_rng(3)
    for rank in range(2, 16):
        for _ in range(5):
            d = W.test2_dims(rng, rank)
            assert len(d) == rank and min(d) >= 2
            assert math.prod(d) <= 2e8
    small = W.test2_dims(np.random.default_rng(1), 9, volmax=4e4)
    assert math.prod(small) <= 4e4
