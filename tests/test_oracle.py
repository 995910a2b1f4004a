"""Pins the CPU oracle (oracle/) against independent references:
libc rand(), hashlib SHA3-256, Python big-int field arithmetic, brute-force
DFT, closed-form Merkle/leaf-chain recomputation, and the restated verifier's
acceptance/rejection behaviour."""
import ctypes
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from virgo_amd import workloads as W

P = (1 << 61) - 1


def to_int(fe):
    return [(int(x["real"]), int(x["img"])) for x in fe]


def from_pairs(pairs):
    a = np.zeros(len(pairs), dtype=O.FE)
    for i, (r, im) in enumerate(pairs):
        a[i]["real"], a[i]["img"] = r, im
    return a


def py_mul(x, y):
    (a, b), (c, d) = x, y
    return ((a * c - b * d) % P, (a * d + b * c) % P)


def rand_fe(rng, n):
    return from_pairs([(int(rng.integers(0, P)), int(rng.integers(0, P))) for _ in range(n)])


def test_rand_matches_libc():
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    want = [libc.rand() for _ in range(5000)]
    got = O.rand_seq(1, 5000).tolist()
    assert got == want
    assert got[:3] == [1804289383, 846930886, 1681692777]


def test_field_ops_bigint():
    rng = np.random.default_rng(5)
    a, b = rand_fe(rng, 400), rand_fe(rng, 400)
    # edge values: 0, 1, p-1 in both components
    edge = from_pairs([(0, 0), (1, 0), (P - 1, P - 1), (0, P - 1), (P - 1, 0), (1, 1)])
    a = np.concatenate([a, edge, edge]); b = np.concatenate([b, edge, edge[::-1]])
    m = to_int(O.fe_mul(a, b)); s = to_int(O.fe_add(a, b)); d = to_int(O.fe_sub(a, b))
    for i, (x, y) in enumerate(zip(to_int(a), to_int(b))):
        assert m[i] == py_mul(x, y)
        assert s[i] == ((x[0] + y[0]) % P, (x[1] + y[1]) % P)
        assert d[i] == ((x[0] - y[0]) % P, (x[1] - y[1]) % P)
    inv = to_int(O.fe_inv(a[:50]))
    for x, y in zip(to_int(a[:50]), inv):
        if x != (0, 0):
            assert py_mul(x, y) == (1, 0)


def test_root_of_unity_order():
    # prime_field.cpp:60-73: order exactly 2^k
    for k in (1, 5, 14, 19, 30, 61):
        w = to_int(O.root_of_unity(k))[0]
        x = w
        for _ in range(k - 1):
            x = py_mul(x, x)
        assert x == (P - 1, 0)          # w^(2^(k-1)) = -1
        assert py_mul(x, x) == (1, 0)


@pytest.mark.parametrize("coef_len,order", [(1, 4), (2, 8), (8, 8), (4, 32), (16, 64), (32, 32)])
def test_fft_brute_force(coef_len, order):
    rng = np.random.default_rng(coef_len * 100 + order)
    c = rand_fe(rng, coef_len)
    lo = order.bit_length() - 1
    w = O.root_of_unity(lo)
    got = to_int(O.fft(c, order, w))
    wi = to_int(w)[0]
    ci = to_int(c)
    pw = [(1, 0)]
    for _ in range(order):
        pw.append(py_mul(pw[-1], wi))
    for k in range(order):
        acc = (0, 0)
        for j in range(coef_len):
            t = py_mul(ci[j], pw[(j * k) % order])
            acc = ((acc[0] + t[0]) % P, (acc[1] + t[1]) % P)
        assert got[k] == acc


def test_ifft_roundtrip_and_subsample():
    rng = np.random.default_rng(7)
    c = rand_fe(rng, 64)
    w = O.root_of_unity(9)
    ev = O.fft(c, 512, w)
    back = O.ifft(ev, 64, w)     # subsampled inverse (RS_polynomial.cpp:183-200)
    assert to_int(back) == to_int(c)


def test_sha3_matches_hashlib():
    rng = np.random.default_rng(3)
    blocks = rng.integers(0, 256, size=(200, 64), dtype=np.uint8)
    blocks[0] = 0
    blocks[1] = 255
    got = O.sha3_64(blocks)
    for i in range(len(blocks)):
        assert bytes(got[i]) == hashlib.sha3_256(bytes(blocks[i])).digest()


def test_merkle_closed_form():
    rng = np.random.default_rng(11)
    for n in (1, 2, 5, 8, 13):
        leaves = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
        tree = O.merkle_tree(leaves)
        pad = len(tree) // 2
        nodes = {}
        for i in range(pad):
            nodes[pad + i] = bytes(leaves[i]) if i < n else hashlib.sha3_256(bytes(64)).digest()
        for k in range(pad - 1, 0, -1):
            nodes[k] = hashlib.sha3_256(nodes[2 * k] + nodes[2 * k + 1]).digest()
        for k in range(1, 2 * pad):
            assert bytes(tree[k]) == nodes[k]


def test_leaf_chain_closed_form():
    # fri.cpp:97-126: t = H((s_j(i), s_j(i+S/2)) || t) over 64 slices, then the mask pair
    rng = np.random.default_rng(2)
    log_S = 3
    S = 1 << log_S
    sl = rand_fe(rng, 64 * S)
    msk = rand_fe(rng, S)
    got = O.leaf_hashes(sl, msk, log_S)
    raw = sl.view(np.uint8).reshape(64, S, 16)
    mr = msk.view(np.uint8).reshape(S, 16)
    for i in range(S // 2):
        t = bytes(32)
        for j in range(64):
            t = hashlib.sha3_256(bytes(raw[j, i]) + bytes(raw[j, i + S // 2]) + t).digest()
        t = hashlib.sha3_256(bytes(mr[i]) + bytes(mr[i + S // 2]) + t).digest()
        assert bytes(got[i]) == t


def test_lde_is_systematic_and_low_degree():
    rng = np.random.default_rng(4)
    vals = rand_fe(rng, 16)
    ext = O.lde_slice(vals, 4, 9)
    assert to_int(ext[::32]) == to_int(vals)       # rate 1/32, systematic
    coef = O.ifft(ext, 512, O.root_of_unity(9))
    assert all(x == (0, 0) for x in to_int(coef[16:]))


def test_fri_fold_halves_degree():
    # folding f(x) = g(x^2) + x h(x^2) with r gives g + r h: degree halves
    rng = np.random.default_rng(8)
    log_S = 6
    S = 1 << log_S
    sl = np.zeros(65 * S, dtype=O.FE)
    for j in range(65):
        c = np.zeros(S, dtype=O.FE)
        c[:8] = rand_fe(rng, 8)
        sl[j * S:(j + 1) * S] = O.fft(c, S, O.root_of_unity(log_S))
    r = rand_fe(rng, 1)
    out = O.fri_fold(sl, log_S, r)
    for j in range(65):
        coef = O.ifft(out[j * S // 2:(j + 1) * S // 2], S // 2, O.root_of_unity(log_S - 1))
        assert all(x == (0, 0) for x in to_int(coef[4:]))


def test_fri_fold1_closed_form():
    """One fold of the codeword of f = sum c_k x^k is the codeword (over the
    squared domain) of sum (c_2k + r c_2k+1) y^k, checked with big-int
    arithmetic on the coefficients; and it equals slice 0 of the 65-slice
    fold of fri.cpp:321-340."""
    rng = np.random.default_rng(12)
    log_S, d = 7, 16
    S = 1 << log_S
    c = np.zeros(S, dtype=O.FE)
    c[:d] = rand_fe(rng, d)
    cw = O.fft(c, S, O.root_of_unity(log_S))
    r = rand_fe(rng, 1)
    got = O.fri_fold1(cw, r)
    ci, ri = to_int(c), to_int(r)[0]
    folded = [((a[0] + py_mul(ri, b)[0]) % P, (a[1] + py_mul(ri, b)[1]) % P) for a, b in zip(ci[0::2], ci[1::2])]
    want = O.fft(from_pairs(folded), S // 2, O.root_of_unity(log_S - 1))
    assert to_int(got) == to_int(want)
    sl = np.zeros(65 * S, dtype=O.FE)
    sl[:S] = cw
    assert to_int(O.fri_fold(sl, log_S, r)[:S // 2]) == to_int(got)


def test_eq_table_bigint():
    rng = np.random.default_rng(9)
    r = rand_fe(rng, 5)
    q = to_int(O.eq_table(r))
    ri = to_int(r)
    for x in range(32):
        acc = (1, 0)
        for i in range(5):
            t = ri[i] if (x >> i) & 1 else ((1 - ri[i][0]) % P, (-ri[i][1]) % P)
            acc = py_mul(acc, t)
        assert q[x] == acc


def int_truncated(x):
    """The input value as the reference stores it: read as long long, kept
    as `int u` (the low 32 bits, two's complement) and widened to unsigned
    long long for field_element(u) (zk_prover.cpp:57-66)."""
    low = x & 0xFFFFFFFF
    signed = low - (1 << 32) if low >= 1 << 31 else low
    return signed % (1 << 64)


def brute_eval(circ):
    """Independent Python evaluation over F_{p^2} of every gate type a
    description may use, by each type's definition (zk_prover.cpp:78-142)."""
    def add(a, b): return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)
    def sub(a, b): return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)
    def const(k): return (k % P, 0)
    one = const(1)
    n = circ["n"]; off = 0
    vals = None
    for li, cnt in enumerate(n):
        ty = circ["ty"][off:off + cnt]; u = circ["u"][off:off + cnt]; v = circ["v"][off:off + cnt]
        off += cnt
        if li == 0:
            vals = [const(int_truncated(int(x))) for x in u]
            continue
        new = []
        for t, a, b in zip(ty.tolist(), u.tolist(), v.tolist()):
            x, y = vals[a], vals[b] if b < len(vals) else None
            if t == 0:
                new.append(add(x, y))
            elif t == 1:
                new.append(py_mul(x, y))
            elif t == 2:
                new.append(const(0))
            elif t == 5:
                s = const(0)
                for k in range(a, b):
                    s = add(s, vals[k])
                new.append(s)
            elif t == 6:
                new.append(sub(one, x))
            elif t == 7:
                new.append(sub(x, y))
            elif t == 8:                            # x + y - 2xy
                new.append(sub(add(x, y), py_mul(const(2), py_mul(x, y))))
            elif t == 9:                            # (1 - x) y
                new.append(py_mul(sub(one, x), y))
            elif t == 10:
                new.append(x)
            elif t == 12:                           # sum of 2^(k-u) x_k over [u, v]
                s = const(0)
                for k in range(a, b + 1):
                    s = add(s, py_mul(const(1 << (k - a)), vals[k]))
                new.append(s)
            elif t == 13:                           # x (1 - x): zero iff x is a bit
                new.append(py_mul(x, sub(one, x)))
            else:
                raise AssertionError("type %d" % t)
        vals = new
    return vals


def test_evaluate_matmul_closed_form():
    c = W.matmul_circuit(4, seed=3)
    out = to_int(O.evaluate(c))
    A = c["u"][:16].reshape(4, 4); B = c["u"][16:32].reshape(4, 4)
    C = (A.astype(object) @ B.astype(object)) % P
    assert [x[0] for x in out] == [int(x) for x in C.ravel()]
    assert out == brute_eval(c)


def test_evaluate_random_circuit():
    c = W.random_circuit(6, 3, seed=4)
    assert to_int(O.evaluate(c)) == brute_eval(c)


@pytest.mark.parametrize("log_width,depth,seed", [(5, 4, 1), (7, 3, 2), (4, 6, 3)])
def test_evaluate_mixed_gate_types(log_width, depth, seed):
    c = W.mixed_circuit(log_width, depth, seed=seed)
    assert set(W.MIXED_TYPES) <= set(c["ty"][c["n"][0]:].tolist())
    assert to_int(O.evaluate(c)) == brute_eval(c)


def test_evaluate_truncates_wide_inputs():
    """Inputs in [2^40, 2^63) keep only their low 32 bits as a signed int;
    about half turn negative and wrap to 2^64 + u before the reduction."""
    c = W.random_circuit(6, 2, seed=8, wide=True)
    ins = c["u"][:64].tolist()
    assert min(ins) >= 1 << 40
    neg = sum(1 for x in ins if x & 0x80000000)
    assert 16 < neg < 48
    assert int_truncated(0x1234_8000_0001) == (1 << 64) - (1 << 31) + 1
    assert to_int(O.evaluate(c)) == brute_eval(c)
    relay = W._pack([(np.full(4, 3, np.int32), np.array([5, (1 << 40) + 7, 0xFFFFFFFF, 1 << 31], np.int64),
                      np.zeros(4, np.int64)),
                     (np.full(4, 10, np.int32), np.arange(4, dtype=np.int64), np.zeros(4, np.int64))])
    # 2^64 mod p = 8: int -1 -> 2^64 - 1 = 7 mod p; int -2^31 -> 8 - 2^31 mod p
    assert [x[0] for x in to_int(O.evaluate(relay))] == [5, 7, 7, (8 - (1 << 31)) % P]


def test_evaluate_bit_gates_on_bits():
    """Closed forms on 0/1 inputs: not, xor, naab and the bit test are the
    boolean functions they name; the power sum reads the bits as a number."""
    bits = [0, 1, 1, 0, 1, 1, 1, 0]
    inp = (np.full(8, 3, np.int32), np.array(bits, np.int64), np.zeros(8, np.int64))
    gates = [(6, 1, 0), (8, 1, 2), (8, 0, 1), (9, 0, 1), (9, 1, 2), (13, 2, 2), (12, 1, 5), (7, 0, 4)]
    lay = (np.array([g[0] for g in gates], np.int32), np.array([g[1] for g in gates], np.int64),
           np.array([g[2] for g in gates], np.int64))
    c = W._pack([inp, lay])
    out = [x[0] for x in to_int(O.evaluate(c))]
    assert out == [0, 0, 1, 1, 0, 0, 0b11011, P - 1]


@pytest.mark.parametrize("seed", [1, 2])
def test_prove_verify_mixed_gate_types(seed):
    c = W.mixed_circuit(6, 4, seed=seed)
    tr, _ = O.prove(c)
    ok, msg = O.verify(c, tr)
    assert ok, msg
    # a sum range one entry longer changes the public predicate: rejected
    c2 = {k: a.copy() for k, a in c.items()}
    base = int(c["n"][0])
    lay = slice(base, base + int(c["n"][1]))
    g = base + int(np.flatnonzero((c["ty"][lay] == 12) & (c["v"][lay] > c["u"][lay]))[0])
    c2["v"][g] -= 1
    assert not O.verify(c2, tr)[0]


@pytest.mark.parametrize("ty,u,v,msg", [(12, 0, 60, "more than 60"), (11, 0, 0, "unknown gate type"),
                                         (13, 1, 2, "u == v"), (5, 3, 3, "sum range")])
def test_loader_rejects_bad_gates(ty, u, v, msg):
    c = W.mixed_circuit(7, 2, seed=5)
    g = int(c["n"][0]) + 3
    c["ty"][g], c["u"][g], c["v"][g] = ty, u, v
    with pytest.raises(Exception, match=msg):
        O.prove(c)


@pytest.fixture(scope="module")
def small_proof():
    c = W.matmul_circuit(8, seed=2)
    tr, _ = O.prove(c)
    return c, tr


def test_prove_verify_accepts(small_proof):
    c, tr = small_proof
    ok, msg = O.verify(c, tr)
    assert ok, msg


@pytest.mark.parametrize("where", [40, 200, 1000, 5000, -40000, -4000, -100])
def test_verify_rejects_mutation(small_proof, where):
    c, tr = small_proof
    b = bytearray(tr)
    i = where % len(b)
    # stay inside a field element's low bytes so the value stays canonical
    b[i] ^= 0x01
    ok, _ = O.verify(c, bytes(b))
    assert not ok


def test_verify_rejects_non_canonical(small_proof):
    """A field element with a component >= p is malformed, also where it is
    congruent to the honest value (V_res + p in its real part)."""
    c, tr = small_proof
    depth = int.from_bytes(tr[4:8], "little")
    hdr = 12 + 4 * depth                             # magic, depth, input bits, bit lengths
    real = int.from_bytes(tr[hdr + 8:hdr + 16], "little")
    b = bytearray(tr)
    b[hdr + 8:hdr + 16] = (real + P).to_bytes(8, "little")
    ok, msg = O.verify(c, bytes(b))
    assert not ok and "canonical" in msg
    # the last element (in the fft_gkr section) with its top bit set
    b = bytearray(tr)
    b[-1] ^= 0x80
    assert not O.verify(c, bytes(b))[0]


def test_verify_rejects_wrong_circuit(small_proof):
    c, tr = small_proof
    c2 = {k: v.copy() for k, v in c.items()}
    # input values are private (committed), so change public wiring instead
    first_gate = int(c["n"][0])
    c2["u"][first_gate] = (c2["u"][first_gate] + 1) % int(c["n"][0])
    ok, _ = O.verify(c2, tr)
    assert not ok
