"""CPU oracle for the LibreTT permutation -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or as the
timed host baseline). The library never calls it.

Two restatements live here:

* ``liboracle.so`` (``oracle.c``): nested loops in C, one output-order
  odometer (plus an OpenMP split of it), one loop written like
  ``tests/example.cpp:175-212`` and one written like
  ``src/TensorTester.cpp:201-220``.
* ``transpose_np``: the same map through ``numpy.transpose`` on a
  column-major view, a library routine that pins the closed form
  independently of the C loops.

Semantics (reference ``tests/example.cpp:175-212``): the input is
column-major with dims ``dim[0..r-1]``; the output is column-major with dims
``dim[perm[0]], dim[perm[1]], ...`` and
``out[sum_r (prod_{s<r} dim[perm[s]]) * i[perm[r]]] = in[sum_r (prod_{s<r} dim[s]) * i[r]]``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ELEM_DTYPES = {4: np.float32, 8: np.float64, 16: np.complex128}


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        for name in ("oracle_transpose", "oracle_transpose_inorder", "oracle_transpose_decode"):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
            f.restype = ctypes.c_int
        L.oracle_transpose_omp.argtypes = [ctypes.c_int, ip, ip, ctypes.c_size_t,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_transpose_omp.restype = ctypes.c_int
        L.oracle_fill_uniform.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int,
                                          ctypes.c_uint64]
        L.oracle_fill_uniform.restype = ctypes.c_int
        L.oracle_max_threads.argtypes = []
        L.oracle_max_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _iarr(v):
    return (ctypes.c_int * len(v))(*[int(x) for x in v])


def _as_bytes_array(a):
    a = np.ascontiguousarray(a)
    return a


def transpose(dim, perm, elem, inp, variant="odometer", nthreads=0):
    """Permute ``inp`` (a flat numpy array of vol elements of ``elem`` bytes).

    ``variant`` selects the C loop: ``odometer`` (output order), ``omp``
    (output order, OpenMP), ``inorder`` (tests/example.cpp style) or
    ``decode`` (src/TensorTester.cpp style).
    """
    inp = _as_bytes_array(inp)
    vol = int(np.prod(dim, dtype=np.int64))
    assert inp.nbytes == vol * elem, (inp.nbytes, vol, elem)
    out = np.empty_like(inp)
    L = lib()
    args = (len(dim), _iarr(dim), _iarr(perm), elem, inp.ctypes.data, out.ctypes.data)
    if variant == "odometer":
        rc = L.oracle_transpose(*args)
    elif variant == "omp":
        rc = L.oracle_transpose_omp(*args, int(nthreads))
    elif variant == "inorder":
        rc = L.oracle_transpose_inorder(*args)
    elif variant == "decode":
        rc = L.oracle_transpose_decode(*args)
    else:
        raise ValueError(variant)
    if rc != 0:
        raise ValueError("oracle rejected arguments dim=%r perm=%r elem=%r" % (dim, perm, elem))
    return out


def transpose_np(dim, perm, inp):
    """The same permutation through numpy.transpose on a Fortran-order view."""
    x = np.asarray(inp).reshape(tuple(dim), order="F")
    return np.transpose(x, tuple(perm)).reshape(-1, order="F")


def fill_uniform(nelem, elem, seed):
    """Seeded values in [-1,1) as a flat array of the element's numpy type."""
    out = np.empty(nelem, dtype=ELEM_DTYPES[elem])
    rc = lib().oracle_fill_uniform(out.ctypes.data, nelem, elem, seed)
    if rc != 0:
        raise ValueError(elem)
    return out


def max_threads():
    return lib().oracle_max_threads()
