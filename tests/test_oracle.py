"""Pins the CPU oracle before anything is checked against it.

The oracle's C loops are compared with
  * hand-worked answers (a 2x3 matrix transpose, a 2x3x2 rank-3 case);
  * numpy.transpose on a column-major view (a library routine);
  * the input-order scatter of tests/example.cpp:175-212 and the per-output
    decode of src/TensorTester.cpp:201-220 (two other forms of the
    reference's own formula);
  * invariants: perm followed by its inverse is the identity, the identity
    perm is a copy, the output is a permutation of the input's elements.
A deliberately wrong oracle (inverse permutation) is shown to fail the
numpy pin, so the pin can tell a transposed operand apart.
"""
import itertools

import numpy as np
import pytest

import oracle
from librett_amd import data

RNG = np.random.default_rng(20261202)


def iota(vol, elem):
    dt = oracle.ELEM_DTYPES[elem]
    if elem == 16:
        return (np.arange(vol) + 1j * np.arange(vol)).astype(dt)
    return np.arange(vol).astype(dt)


def test_hand_worked_matrix():
    # 2x3 column-major matrix a[i,j] at i + 2j, transposed to 3x2.
    x = np.arange(6, dtype=np.float64)
    out = oracle.transpose([2, 3], [1, 0], 8, x)
    assert out.tolist() == [0, 2, 4, 1, 3, 5]


def test_hand_worked_rank3():
    # dims (2,3,2), perm (2,0,1): output dims (2,2,3),
    # out[k + 2 i + 4 j] = in[i + 2 j + 6 k].
    x = np.arange(12, dtype=np.float32)
    out = oracle.transpose([2, 3, 2], [2, 0, 1], 4, x)
    exp = np.empty(12, dtype=np.float32)
    for i in range(2):
        for j in range(3):
            for k in range(2):
                exp[k + 2 * i + 4 * j] = x[i + 2 * j + 6 * k]
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("variant", ["odometer", "omp", "inorder", "decode"])
@pytest.mark.parametrize("elem", [4, 8, 16])
def test_matches_numpy_all_perms(variant, elem):
    for rank in range(1, 6):
        dim = [2 + r * (9 - 2) // max(1, rank) for r in range(rank)]
        vol = int(np.prod(dim))
        x = data.uniform(vol, elem, 11 + rank)
        for perm in itertools.permutations(range(rank)):
            out = oracle.transpose(dim, list(perm), elem, x, variant=variant)
            ref = oracle.transpose_np(dim, perm, x)
            assert np.array_equal(out.view(np.uint8), ref.view(np.uint8)), (dim, perm)


@pytest.mark.parametrize("elem", [4, 8, 16])
def test_random_shapes_match_numpy(elem):
    for _ in range(40):
        rank = int(RNG.integers(1, 8))
        dim = [int(d) for d in RNG.integers(1, 7, size=rank)]
        perm = [int(p) for p in RNG.permutation(rank)]
        x = data.uniform(int(np.prod(dim)), elem, int(RNG.integers(1 << 30)))
        out = oracle.transpose(dim, perm, elem, x)
        assert np.array_equal(out.view(np.uint8), oracle.transpose_np(dim, perm, x).view(np.uint8))


def test_example_cpp_known_answer():
    # tests/example.cpp:88-102: dims {3,3,3,2,4}, perm {1,3,0,4,2}, inp[i]=i,
    # check out[tindex] == inp[index] with tindex = sum cba[r]*ijk[perm[r]].
    dim, perm = [3, 3, 3, 2, 4], [1, 3, 0, 4, 2]
    vol = int(np.prod(dim))
    for elem in (4, 8, 16):
        x = iota(vol, elem)
        out = oracle.transpose(dim, perm, elem, x)
        abc = np.cumprod([1] + dim[:-1])
        cba = np.cumprod([1] + [dim[p] for p in perm[:-1]])
        for index in range(vol):
            ijk = [(index // abc[r]) % dim[r] for r in range(5)]
            t = sum(cba[r] * ijk[perm[r]] for r in range(5))
            assert out[t] == x[index]


@pytest.mark.parametrize("elem", [4, 8, 16])
def test_inverse_roundtrip_and_identity(elem):
    for _ in range(20):
        rank = int(RNG.integers(1, 7))
        dim = [int(d) for d in RNG.integers(1, 8, size=rank)]
        perm = [int(p) for p in RNG.permutation(rank)]
        inv = [0] * rank
        for i, p in enumerate(perm):
            inv[p] = i
        x = data.uniform(int(np.prod(dim)), elem, 5)
        y = oracle.transpose(dim, perm, elem, x)
        odim = [dim[p] for p in perm]
        z = oracle.transpose(odim, inv, elem, y)
        assert np.array_equal(z.view(np.uint8), x.view(np.uint8))
        ident = oracle.transpose(dim, list(range(rank)), elem, x)
        assert np.array_equal(ident.view(np.uint8), x.view(np.uint8))
        assert np.array_equal(np.sort(y.view(np.uint8).reshape(-1, elem), axis=0),
                              np.sort(x.view(np.uint8).reshape(-1, elem), axis=0))


def test_wrong_oracle_is_caught():
    # A plausible mistake (using the inverse permutation) must fail the numpy
    # pin for every non-involution permutation.
    dim, perm = [2, 3, 4], [1, 2, 0]
    inv = [2, 0, 1]
    x = data.uniform(24, 8, 3)
    wrong = oracle.transpose([dim[i] for i in range(3)], inv, 8, x)
    assert not np.array_equal(wrong, oracle.transpose_np(dim, perm, x))


def test_rejects_invalid_arguments():
    x = np.zeros(4, dtype=np.float32)
    for dim, perm, elem in [([2, 2], [0, 0], 4), ([2, 2], [0, 2], 4), ([2, 2], [1, 0], 2), ([0, 4], [1, 0], 4)]:
        with pytest.raises((ValueError, AssertionError)):
            oracle.transpose(dim, perm, elem, x)


@pytest.mark.parametrize("elem", [4, 8, 16])
def test_generators_agree(elem):
    # The C oracle's generator and the package's numpy generator are two
    # implementations of the same counter hash.
    a = data.uniform(4097, elem, 987654321)
    b = oracle.fill_uniform(4097, elem, 987654321)
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    f = a.view(np.float32 if elem == 4 else np.float64)
    assert f.min() >= -1.0 and f.max() < 1.0
