"""Index-pattern checker (the reference's TensorTester counterpart) on the
host: the same per-element code the GPU kernel runs, against the oracle."""
import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import workloads as W


def pattern(dim, elem):
    x = np.zeros(int(np.prod(dim)) * elem // 4, dtype=np.uint32)
    L.host_check_fill(x)
    return x


def test_fill_is_word_index():
    x = np.zeros(1000, dtype=np.uint32)
    L.host_check_fill(x)
    assert np.array_equal(x, np.arange(1000, dtype=np.uint32))


@pytest.mark.parametrize("elem", [4, 8, 16])
def test_correct_transposes_pass(elem):
    rng = np.random.default_rng(elem)
    for _ in range(40):
        d, p, e = W.random_problem(rng, 1, 6, 1, 5000, [elem])
        x = pattern(d, e)
        y = oracle.transpose(d, p, e, x.view(oracle.ELEM_DTYPES[e]))
        assert L.host_check_transpose(y, d, p, e) == (0, -1), (d, p, e)


@pytest.mark.parametrize("elem", [4, 8, 16])
def test_planted_errors_are_found(elem):
    rng = np.random.default_rng(100 + elem)
    d, p = [7, 5, 11, 3], [2, 0, 3, 1]
    x = pattern(d, elem)
    y = oracle.transpose(d, p, elem, x.view(oracle.ELEM_DTYPES[elem])).view(np.uint32).copy()
    words = elem // 4
    bad_words = rng.choice(y.size, 9, replace=False)
    y[bad_words] ^= np.uint32(0x80000000)
    bad_elems = sorted(set(int(w) // words for w in bad_words))
    n, first = L.host_check_transpose(y, d, p, elem)
    assert n == len(bad_elems) and first == bad_elems[0]


def test_wrong_permutation_is_found():
    d, p = [4, 6, 5], [1, 2, 0]
    inv = [2, 0, 1]
    x = pattern(d, 8)
    y = oracle.transpose([d[i] for i in inv], p, 8, x.view(np.float64))  # same volume, other map
    n, first = L.host_check_transpose(y, d, p, 8)
    assert n > 0 and first >= 0


def test_checker_rejects_bad_input():
    y = np.zeros(64, dtype=np.uint32)
    for dim, perm, e in [([4, 4], [0, 0], 4), ([4, 4], [0, 2], 4), ([4, 4], [1, 0], 2), ([0, 4], [1, 0], 4)]:
        with pytest.raises(ValueError):
            L.host_check_transpose(y, dim, perm, e)
