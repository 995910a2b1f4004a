"""GPU: the kernel micro-set at full sizes -- forward and inverse FFT over
F_{p^2} at 2^18 ... 2^24 with uniform random coefficients, the rate-1/8
Reed-Solomon extension, the Merkle tree over its 32-byte leaves (value
pairs f(w^j), f(-w^j)) and FRI folding down to a 32-value codeword.

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
