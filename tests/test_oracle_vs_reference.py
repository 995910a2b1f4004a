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


REF_FG = os.path.join(ROOT, "oracle", "_ref", "libref_fftgkr.so")


def _ref_fg():
    if not os.path.exists(REF_FG):
        pytest.skip("oracle/_ref/libref_fftgkr.so not built (needs the reference sources)")
    L = ctypes.CDLL(REF_FG)
    L.ref_fg_run.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    for name in ("ref_fg_quad", "ref_fg_lin"):
        getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_long]
        getattr(L, name).restype = ctypes.c_long
    L.ref_fg_layer.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_long]
    L.ref_fg_layer.restype = ctypes.c_long
    return L


def _records(L, name):
    n = getattr(L, name)(None, 0)
    out = np.zeros(4 * n, dtype=O.FE)
    getattr(L, name)(_p(out), n)
    return out.reshape(n, 4)


def _section_fields(sec, lg):
    """Splits an oracle "VGF1" section into (a_0, round polynomials, layer
    values v_u / v_v) in the order the section stores them."""
    assert sec[:4] == b"VGF1" and int.from_bytes(sec[4:8], "little") == lg
    f = np.frombuffer(sec[8:], dtype=O.FE)
    a0, k = f[0], 1
    polys, vals = [], []

    def rounds(m):
        nonlocal k
        for _ in range(m):
            polys.append(f[k:k + 3])
            k += 3

    def val():
        nonlocal k
        vals.append(f[k])
        k += 1
    rounds(lg + 6); val()                   # addition layer, v_u
    rounds(lg); val()                       # mult layer, v_u
    for _ in range(lg):                     # inverse-DFT layers: v_u, v_v
        rounds(lg); val(); rounds(lg); val()
    assert k == len(f)
    return a0, polys, vals


@pytest.mark.parametrize("lg", [1, 4, 9, 14])
def test_fft_gkr_matches_reference_program(lg):
    """The oracle's fft_gkr against the reference's own program
    (fft_circuit_GKR.cpp compiled from source, main body run on a fresh
    rand() stream): every circuit layer, every round polynomial its prover
    makes (recorded at each quadratic_poly::eval), every v_u / v_v (each
    linear_poly::eval), the claim on the outputs, the verdict and the proof
    size equal the oracle's section."""
    L = _ref_fg()
    res, psize = ctypes.c_int(), ctypes.c_int()
    nl = L.ref_fg_run(lg, ctypes.byref(res), ctypes.byref(psize))
    sec, ok, osize, _ = O.fft_gkr(lg)
    assert res.value == 1 and ok
    assert psize.value == osize
    for idx in range(nl):
        n = L.ref_fg_layer(idx, None, 0)
        ref = np.zeros(n, dtype=O.FE)
        L.ref_fg_layer(idx, _p(ref), n)
        assert np.array_equal(ref, O.fft_gkr_layer(lg, idx)), "layer %d" % idx
    a0, polys, vals = _section_fields(sec, lg)
    quad = _records(L, "ref_fg_quad")
    # each round evaluates its polynomial at 0 and 1 (the check) and at the challenge
    assert len(quad) == 3 * len(polys)
    for k, p in enumerate(polys):
        grp = quad[3 * k:3 * k + 3]
        assert all(np.array_equal(g[:3], grp[0][:3]) for g in grp), "round %d" % k
        assert sorted(int(g[3]["real"]) for g in grp[:2]) == [0, 1] and not grp[0][3]["img"]
        assert np.array_equal(grp[0][:3], p), "round polynomial %d" % k
    # claim on the 64 outputs = p(0) + p(1) of the first round = a + b + 2c
    a, b, c = (O.fe_add(polys[0][i:i + 1], np.zeros(1, O.FE)) for i in range(3))
    assert np.array_equal(O.fe_add(O.fe_add(a, b), O.fe_add(c, c)), a0.reshape(1))
    lin = _records(L, "ref_fg_lin")
    assert len(lin) == len(vals)
    for k, v in enumerate(vals):
        assert lin[k][3] == v, "layer value %d" % k
