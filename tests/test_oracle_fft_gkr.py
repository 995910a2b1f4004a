"""Pins the oracle's fft_gkr restatement (fft_circuit_GKR.cpp:22-845).

The circuit layers are checked against their closed forms with Python
big-int arithmetic: layer lg is the product table of the first lg challenges,
layer 2lg+1 is its natural-order inverse DFT, the 64 outputs are that
polynomial at the next 64 draws, and the first message is the multilinear
extension of the outputs at r_0.  The sumcheck messages are pinned by the
program's own verifier equations (round sums and the closed forms that end
each layer), which must accept honest sections and reject any changed one.
"""
import numpy as np
import pytest

from oracle import oracle as O

P = (1 << 61) - 1
Q = P * P - 1          # order of the multiplicative group of F_{p^2}


def fmul(x, y):
    (a, b), (c, d) = x, y
    return ((a * c - b * d) % P, (a * d + b * c) % P)


def fadd(x, y):
    return ((x[0] + y[0]) % P, (x[1] + y[1]) % P)


def fsub(x, y):
    return ((x[0] - y[0]) % P, (x[1] - y[1]) % P)


def fpow(x, e):
    r = (1, 0)
    while e:
        if e & 1:
            r = fmul(r, x)
        x = fmul(x, x)
        e >>= 1
    return r


def pairs(a):
    return [(int(x["real"]), int(x["img"])) for x in a]


class Draws:
    """prime_field::random on glibc rand() seed 1: real first, then img."""

    def __init__(self, n):
        self.raw = [int(x) for x in O.rand_seq(1, 2 * n)]
        self.k = 0

    def fe(self):
        r, i = self.raw[self.k], self.raw[self.k + 1]
        self.k += 2
        return (r % P, i % P)


def root(lg):
    w = O.root_of_unity(lg)
    w = np.atleast_1d(w)[0]
    return (int(w["real"]), int(w["img"]))


@pytest.mark.parametrize("lg", [3, 5])
def test_circuit_layers_closed_form(lg):
    N = 1 << lg
    d = Draws(lg + 64 + 2 * (lg + 10))
    r = [d.fe() for _ in range(lg)]
    # layer lg: value[x] = prod_i (bit (lg-1-i) of x ? 1 - r_i : r_i)
    e = []
    for x in range(N):
        v = (1, 0)
        for i in range(lg):
            bit = (x >> (lg - 1 - i)) & 1
            v = fmul(v, fsub((1, 0), r[i]) if bit else r[i])
        e.append(v)
    assert pairs(O.fft_gkr_layer(lg, lg)) == e
    # layer 2lg+1: natural-order inverse DFT of layer lg
    winv = fpow(root(lg), Q - 1)
    ninv = fpow((N, 0), Q - 1)
    coef = []
    for k in range(N):
        s = (0, 0)
        for x in range(N):
            s = fadd(s, fmul(e[x], fpow(winv, x * k)))
        coef.append(fmul(s, ninv))
    assert pairs(O.fft_gkr_layer(lg, 2 * lg + 1)) == coef
    # outputs: the polynomial at 64 fresh draws
    pts = [d.fe() for _ in range(64)]
    outs = []
    for x in pts:
        s, xp = (0, 0), (1, 0)
        for c in coef:
            s = fadd(s, fmul(c, xp))
            xp = fmul(xp, x)
        outs.append(s)
    assert pairs(O.fft_gkr_layer(lg, 2 * lg + 3)) == outs
    # first message: MLE of the outputs at r_0 (bit 0 = lowest index bit)
    r0 = [d.fe() for _ in range(lg + 10)]
    t = list(outs)
    for i in range(6):
        t = [fadd(fmul(t[2 * j], fsub((1, 0), r0[i])), fmul(t[2 * j + 1], r0[i])) for j in range(len(t) // 2)]
    sec, ok, _, _ = O.fft_gkr(lg)
    a0 = np.frombuffer(sec[8:24], dtype=O.FE)
    assert pairs(a0) == t and ok


@pytest.mark.parametrize("lg", [1, 2, 4, 9, 11])
def test_accepts_and_proof_size(lg):
    sec, ok, psize, _ = O.fft_gkr(lg)
    assert ok
    rounds = (lg + 6) + lg + 2 * lg * lg
    assert len(sec) == 8 + 16 + 48 * rounds + 16 * (2 + 2 * lg)
    assert psize == 48 * (rounds + lg * (lg + 1) // 2)


@pytest.mark.parametrize("where", [8, 30, 100, 700, -500, -60, -20])
def test_rejects_changed_message(where):
    lg = 6
    sec, ok, _, _ = O.fft_gkr(lg)
    assert ok and O.fft_gkr_verify(lg, sec)
    b = bytearray(sec)
    b[where % len(b)] ^= 0x01
    assert not O.fft_gkr_verify(lg, bytes(b))
    assert not O.fft_gkr_verify(lg, sec[:-16])
    assert not O.fft_gkr_verify(lg + 1, sec)
