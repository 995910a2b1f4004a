"""GPU: the kernel micro-set at full sizes -- forward and inverse FFT over
F_{p^2} at 2^18 ... 2^24 with uniform random coefficients, the rate-1/8
Reed-Solomon extension, the Merkle tree over its 32-byte leaves (value
pairs f(w^j), f(-w^j)) and FRI folding down to a 32-value codeword.

The micro-set is also run as one call (vg_microset: device-resident,
timed per stage, SHA3-256 or SHA-256 trees, FRI commit roots of every
level) and checked against the single operations.  SHA-256 is checked
against hashlib (it has no counterpart in the reference).

Bit-exact against the oracle where it finishes in seconds (FFT up to 2^20,
extension, tree and first folds at 2^18); at every size through properties
that do not depend on size: FFT/iFFT round trip, the extension restricted
to the subgroup equals the plain transform, every tree node is SHA3-256 of
its children (hashlib), and the final folded codeword equals the closed
form (the coefficient fold c'_k = c_2k + r c_2k+1 applied L-5 times with
the oracle's field ops, then evaluated), which also has degree < 4."""
import hashlib

import numpy as np
import pytest

import virgo_amd as V
from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = (1 << 61) - 1
RATE_LOG = 3          # rate 1/8


def rand_fe(rng, n):
    a = np.zeros(n, dtype=V.FE)
    a["real"] = rng.integers(0, P, n, dtype=np.uint64)
    a["img"] = rng.integers(0, P, n, dtype=np.uint64)
    return a


@pytest.mark.parametrize("log_n", [18, 20, 22, 24])
def test_kernel_micro_set(log_n):
    print("kernel micro-set, log_n", log_n, flush=True)
    N, L = 1 << log_n, log_n + RATE_LOG
    M = 1 << L
    rng = np.random.default_rng(100 + log_n)
    c = rand_fe(rng, N)
    w, wl = O.root_of_unity(log_n), O.root_of_unity(L)
    # forward + inverse transform of size N
    f = V.fft(c, N, w)
    assert np.array_equal(V.ifft(f, N, w), c)
    if log_n <= 20:
        assert np.array_equal(f, O.fft(c, N, w))
    # rate-1/8 extension: the transform of c on 8N points
    E = V.fft(c, M, wl)
    assert np.array_equal(E[::1 << RATE_LOG], f)
    if log_n == 18:
        assert np.array_equal(E, O.fft(c, M, wl))
    # Merkle tree over the 32-byte leaves (E[j], E[j + M/2])
    leaves = np.ascontiguousarray(np.stack([E[:M // 2], E[M // 2:]], axis=1)).view(np.uint8).reshape(-1, 32)
    tree = V.merkle_tree(leaves)
    pad = len(tree) // 2
    assert pad == M // 2
    if log_n == 18:
        assert np.array_equal(tree, O.merkle_tree(leaves))
    else:
        assert np.array_equal(tree[pad:], leaves)
        for k in list(rng.integers(1, pad, 200)) + [1, 2, 3, pad - 1]:
            assert bytes(tree[k]) == hashlib.sha3_256(bytes(tree[2 * k]) + bytes(tree[2 * k + 1])).digest()
    del leaves, tree
    # FRI folding down to 32 values
    rs = rand_fe(rng, L - 5)
    cw = E
    for k in range(L - 5):
        nxt = V.fri_fold(cw, rs[k:k + 1])
        if log_n == 18 and k < 3:
            assert np.array_equal(nxt, O.fri_fold1(cw, rs[k:k + 1]))
        cw = nxt
    assert len(cw) == 32
    cf = np.zeros(M, dtype=O.FE)
    cf[:N] = c
    for k in range(L - 5):
        cf = O.fe_add(cf[0::2], O.fe_mul(np.repeat(rs[k:k + 1], len(cf) // 2), cf[1::2]))
    assert not cf[4:].view(np.uint64).any()          # degree < 32 / 8
    assert np.array_equal(cw, O.fft(cf, 32, O.root_of_unity(5)))


def test_sha256_matches_hashlib():
    rng = np.random.default_rng(7)
    blocks = rng.integers(0, 256, (1000, 64), dtype=np.uint8)
    blocks[0] = 0
    out = V.sha256_64(blocks)
    for k in range(len(blocks)):
        assert bytes(out[k]) == hashlib.sha256(bytes(blocks[k])).digest()


@pytest.mark.parametrize("n_leaf", [1, 5, 1000, 4096, 70000])
def test_merkle_sha256_matches_hashlib(n_leaf):
    """Heap layout of vg_merkle_tree with SHA-256 nodes; pad leaves are
    SHA-256 of 64 zero bytes."""
    rng = np.random.default_rng(n_leaf)
    leaves = rng.integers(0, 256, (n_leaf, 32), dtype=np.uint8)
    tree = V.merkle_tree(leaves, hash="sha256")
    pad = len(tree) // 2
    level = [bytes(x) for x in leaves] + [hashlib.sha256(bytes(64)).digest()] * (pad - n_leaf)
    assert [bytes(x) for x in tree[pad:]] == level
    while len(level) > 1:
        level = [hashlib.sha256(level[2 * i] + level[2 * i + 1]).digest() for i in range(len(level) // 2)]
    assert bytes(tree[1]) == level[0]


def _py_root(leaves, h):
    level = [bytes(x) for x in leaves]
    while len(level) > 1:
        level = [h(level[2 * i] + level[2 * i + 1]).digest() for i in range(len(level) // 2)]
    return level[0]


@pytest.mark.parametrize("log_n,hash,rate_log", [(18, "sha3", RATE_LOG), (18, "sha256", RATE_LOG), (20, "sha256", RATE_LOG),
                                                  (22, "sha256", RATE_LOG), (24, "sha3", RATE_LOG), (16, "sha3", 5)])
def test_microset_stages(log_n, hash, rate_log):
    """vg_microset (one call, device-resident, timed per stage) against an
    independent route: every FRI level's codeword is the transform of the
    coefficients folded in the coefficient domain (c'_k = c_2k + r c_2k+1),
    made by the single-operation FFT; each level's root is the
    single-operation tree of its pairs (and, at 2^18, a hashlib tree).  The
    last case runs at the reference's rate 1/32."""
    print("microset", log_n, hash, rate_log, flush=True)
    N, L = 1 << log_n, log_n + rate_log
    M = 1 << L
    rng = np.random.default_rng(300 + log_n)
    c = rand_fe(rng, N)
    rs = rand_fe(rng, L - 5)
    roots, final, ok, ms = V.microset(c, rate_log, rs, hash=hash, reps=2)
    assert ok
    assert all(v > 0 for v in ms.values()), ms
    assert roots.shape == (L - 5, 32)
    cf = np.zeros(M, dtype=O.FE)
    cf[:N] = c
    for k in range(L - 5):
        n_k = M >> k
        cw = V.fft(cf[:n_k], n_k, O.root_of_unity(L - k))
        leaves = np.ascontiguousarray(np.stack([cw[:n_k // 2], cw[n_k // 2:]], axis=1)).view(np.uint8).reshape(-1, 32)
        want = V.merkle_tree(leaves, hash=hash)[1].tobytes()
        assert roots[k].tobytes() == want, "level %d" % k
        if log_n == 18 and k in (0, L - 7):
            assert want == _py_root(leaves, hashlib.sha3_256 if hash == "sha3" else hashlib.sha256)
        del cw, leaves
        cf = O.fe_add(cf[0::2], O.fe_mul(np.repeat(rs[k:k + 1], n_k // 2), cf[1::2]))
    assert np.array_equal(final, O.fft(cf, 32, O.root_of_unity(5)))


def test_valu_issue_rates_match_the_roofline_model():
    """The leaf kernel's roofline prices three-input bitwise operations at 2
    cycles and funnel shifts at 4 per wave64: the device's measured issue
    rates (vg_valu_rates) must show that split."""
    r = V.valu_rates()
    assert 1000 < r["clock_mhz"] < 3000
    assert 0.3 < r["bitwise_per_cycle"] <= 0.5 and 0.15 < r["shift_per_cycle"] <= 0.25
    assert r["bitwise_per_cycle"] > 1.5 * r["shift_per_cycle"]
