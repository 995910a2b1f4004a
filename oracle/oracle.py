"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: a single-threaded C restatement of the reference
Virgo prover/verifier used as the bit-exact checker for the HIP product and as
the CPU baseline timer.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FE = np.dtype([("img", "<u8"), ("real", "<u8")])   # prime_field.h:93 order
P61 = (1 << 61) - 1


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.vgo_prove.argtypes = [vp, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t), vp]
        L.vgo_verify.argtypes = [vp, vp, ctypes.c_size_t]
        L.vgo_evaluate.argtypes = [vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]
        L.vgo_last_error.restype = ctypes.c_char_p
        L.vgo_free.argtypes = [vp]
        for name in ("vgo_fe_mul", "vgo_fe_add", "vgo_fe_sub"):
            getattr(L, name).argtypes = [vp, vp, vp, ctypes.c_size_t]
        L.vgo_fe_inv.argtypes = [vp, vp, ctypes.c_size_t]
        L.vgo_root_of_unity.argtypes = [ctypes.c_int, vp]
        L.vgo_rand_seq.argtypes = [ctypes.c_uint32, vp, ctypes.c_size_t]
        L.vgo_fft.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.vgo_ifft.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.vgo_sha3_64.argtypes = [vp, vp, ctypes.c_size_t]
        L.vgo_merkle_tree.argtypes = [vp, ctypes.c_int, vp]
        L.vgo_lde_slice.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        L.vgo_leaf_hashes.argtypes = [vp, vp, ctypes.c_int, vp]
        L.vgo_fri_fold.argtypes = [vp, ctypes.c_int, vp, vp]
        L.vgo_fri_fold1.argtypes = [vp, ctypes.c_int, vp, vp]
        L.vgo_eq_table.argtypes = [vp, ctypes.c_int, vp]
        L.vgo_fft_gkr.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t),
                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_double)]
        L.vgo_fft_gkr_verify.argtypes = [ctypes.c_int, vp, ctypes.c_size_t]
        L.vgo_fft_gkr_layer.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_size_t]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _Desc(ctypes.Structure):
    _fields_ = [("d", ctypes.c_int), ("n", ctypes.c_void_p), ("ty", ctypes.c_void_p),
                ("u", ctypes.c_void_p), ("v", ctypes.c_void_p)]


def _desc(circ):
    n = np.ascontiguousarray(circ["n"], dtype=np.int64)
    ty = np.ascontiguousarray(circ["ty"], dtype=np.int32)
    u = np.ascontiguousarray(circ["u"], dtype=np.int64)
    v = np.ascontiguousarray(circ["v"], dtype=np.int64)
    keep = (n, ty, u, v)
    d = _Desc(len(n), _p(n), _p(ty), _p(u), _p(v))
    return d, keep


def prove(circ):
    """Returns (transcript bytes, times[evaluate, gkr, commit, queries, fft_gkr])."""
    L = lib()
    d, keep = _desc(circ)
    out = ctypes.c_void_p()
    ln = ctypes.c_size_t()
    times = np.zeros(5, dtype=np.float64)
    rc = L.vgo_prove(ctypes.byref(d), ctypes.byref(out), ctypes.byref(ln), _p(times))
    if rc != 0:
        raise RuntimeError("oracle prove failed (%d): %s" % (rc, L.vgo_last_error().decode()))
    data = ctypes.string_at(out, ln.value)
    L.vgo_free(out)
    return data, times


def verify(circ, transcript):
    L = lib()
    d, keep = _desc(circ)
    buf = np.frombuffer(transcript, dtype=np.uint8)
    rc = L.vgo_verify(ctypes.byref(d), _p(buf), len(transcript))
    return rc == 1, L.vgo_last_error().decode()


def fft_gkr(lg):
    """fft_gkr prover (fresh seed-1 stream): returns (transcript section,
    accepted, proof_size, prover seconds)."""
    L = lib()
    out = ctypes.c_void_p()
    ln = ctypes.c_size_t()
    acc, ps, sec = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    rc = L.vgo_fft_gkr(lg, ctypes.byref(out), ctypes.byref(ln), ctypes.byref(acc), ctypes.byref(ps), ctypes.byref(sec))
    if rc != 0:
        raise RuntimeError(L.vgo_last_error().decode())
    data = ctypes.string_at(out, ln.value)
    L.vgo_free(out)
    return data, bool(acc.value), ps.value, sec.value


def fft_gkr_verify(lg, section):
    buf = np.frombuffer(section, dtype=np.uint8)
    return lib().vgo_fft_gkr_verify(lg, _p(buf), len(section)) == 1


def fft_gkr_layer(lg, idx):
    cap = 64 << lg
    out = np.zeros(cap, dtype=FE)
    n = lib().vgo_fft_gkr_layer(lg, idx, _p(out), cap)
    if n < 0:
        raise RuntimeError(lib().vgo_last_error().decode())
    return out[:n].copy()


def evaluate(circ):
    L = lib()
    d, keep = _desc(circ)
    cap = 1 << 26
    out = np.zeros(cap, dtype=FE)
    bits = ctypes.c_int()
    rc = L.vgo_evaluate(ctypes.byref(d), _p(out), cap, ctypes.byref(bits))
    if rc != 0:
        raise RuntimeError(L.vgo_last_error().decode())
    return out[: 1 << bits.value].copy()


def _binop(name, a, b):
    a = np.ascontiguousarray(a, dtype=FE)
    b = np.ascontiguousarray(b, dtype=FE)
    o = np.zeros_like(a)
    getattr(lib(), name)(_p(a), _p(b), _p(o), len(a))
    return o


def fe_mul(a, b):
    return _binop("vgo_fe_mul", a, b)


def fe_add(a, b):
    return _binop("vgo_fe_add", a, b)


def fe_sub(a, b):
    return _binop("vgo_fe_sub", a, b)


def fe_inv(a):
    a = np.ascontiguousarray(a, dtype=FE)
    o = np.zeros_like(a)
    lib().vgo_fe_inv(_p(a), _p(o), len(a))
    return o


def root_of_unity(log_order):
    o = np.zeros(1, dtype=FE)
    lib().vgo_root_of_unity(log_order, _p(o))
    return o


def rand_seq(seed, n):
    o = np.zeros(n, dtype=np.int32)
    lib().vgo_rand_seq(seed, _p(o), n)
    return o


def fft(coef, order, rou):
    coef = np.ascontiguousarray(coef, dtype=FE)
    o = np.zeros(order, dtype=FE)
    lib().vgo_fft(_p(coef), len(coef), order, _p(np.ascontiguousarray(rou, dtype=FE)), _p(o))
    return o


def ifft(evals, coef_len, rou):
    evals = np.ascontiguousarray(evals, dtype=FE)
    o = np.zeros(coef_len, dtype=FE)
    lib().vgo_ifft(_p(evals), coef_len, len(evals), _p(np.ascontiguousarray(rou, dtype=FE)), _p(o))
    return o


def sha3_64(blocks):
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1, 64)
    o = np.zeros((len(blocks), 32), dtype=np.uint8)
    lib().vgo_sha3_64(_p(blocks), _p(o), len(blocks))
    return o


def merkle_tree(leaves):
    leaves = np.ascontiguousarray(leaves, dtype=np.uint8).reshape(-1, 32)
    n = len(leaves)
    pad = 1
    while pad < n:
        pad *= 2
    tree = np.zeros((2 * pad, 32), dtype=np.uint8)
    lib().vgo_merkle_tree(_p(leaves), n, _p(tree))
    return tree


def lde_slice(vals, log_real, log_slice):
    vals = np.ascontiguousarray(vals, dtype=FE)
    o = np.zeros(1 << log_slice, dtype=FE)
    lib().vgo_lde_slice(_p(vals), log_real, log_slice, _p(o))
    return o


def leaf_hashes(slices, mask, log_S):
    slices = np.ascontiguousarray(slices, dtype=FE)
    mask = np.ascontiguousarray(mask, dtype=FE)
    o = np.zeros(((1 << log_S) // 2, 32), dtype=np.uint8)
    lib().vgo_leaf_hashes(_p(slices), _p(mask), log_S, _p(o))
    return o


def fri_fold(slices, log_S, r):
    slices = np.ascontiguousarray(slices, dtype=FE)
    o = np.zeros(65 * ((1 << log_S) // 2), dtype=FE)
    lib().vgo_fri_fold(_p(slices), log_S, _p(np.ascontiguousarray(r, dtype=FE)), _p(o))
    return o


def fri_fold1(cw, r):
    """One FRI fold of a single natural-order codeword (fri.cpp:321-340)."""
    cw = np.ascontiguousarray(cw, dtype=FE)
    o = np.zeros(len(cw) // 2, dtype=FE)
    lib().vgo_fri_fold1(_p(cw), len(cw).bit_length() - 1, _p(np.ascontiguousarray(r, dtype=FE).reshape(1)), _p(o))
    return o


def eq_table(r):
    r = np.ascontiguousarray(r, dtype=FE)
    o = np.zeros(1 << len(r), dtype=FE)
    lib().vgo_eq_table(_p(r), len(r), _p(o))
    return o
