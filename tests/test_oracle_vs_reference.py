"""Pins the oracle's F_{p^2} arithmetic against the reference's own
prime_field.cpp / polynomial.cpp, compiled from source into oracle/_ref by
oracle/ref_field.mk (skipped where the reference tree is absent)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_field.so")
P = (1 << 61) - 1


def _ref():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libref_field.so not built (needs the reference sources)")
    L = ctypes.CDLL(REF_SO)
    L.ref_init()
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _rand(rng, n):
    a = np.zeros(n, dtype=O.FE)
    a["real"] = rng.integers(0, P, n, dtype=np.uint64)
    a["img"] = rng.integers(0, P, n, dtype=np.uint64)
    # canonical edge values
    edge = np.array([(0, 0), (0, 1), (P - 1, P - 1), (1, P - 1), (P - 1, 0)], dtype=O.FE)
    a[: len(edge)] = edge
    return a


def test_mul_add_sub_inv_match_reference():
    L = _ref()
    rng = np.random.default_rng(21)
    n = 20000
    a, b = _rand(rng, n), _rand(rng, n)
    b[:5] = b[:5][::-1]
    for name, fn in (("ref_mul", O.fe_mul), ("ref_add", O.fe_add), ("ref_sub", O.fe_sub)):
        o = np.zeros(n, dtype=O.FE)
        getattr(L, name)(_p(a), _p(b), _p(o), ctypes.c_long(n))
        assert np.array_equal(o, fn(a, b)), name
    o = np.zeros(n, dtype=O.FE)
    L.ref_packed_mul(_p(a), _p(b), _p(o), ctypes.c_long(n))
    assert np.array_equal(o, O.fe_mul(a, b))
    a1 = a[5:300]
    o = np.zeros(len(a1), dtype=O.FE)
    L.ref_inv(_p(a1), _p(o), ctypes.c_long(len(a1)))
    assert np.array_equal(o, O.fe_inv(a1))


def test_roots_match_reference():
    L = _ref()
    for k in range(0, 62):
        o = np.zeros(1, dtype=O.FE)
        L.ref_root(k, _p(o))
        assert np.array_equal(o, O.root_of_unity(k)), k
