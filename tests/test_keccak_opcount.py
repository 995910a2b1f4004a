"""Derives the algorithmic operation count per Keccak-f[1600] that the bench's
roofline divides by, and pins the bench's constant to it.

Keccak-f[1600] on 32-bit halves (FIPS 202 round: theta as five column
parities, five 1-bit rotations and one three-input XOR per half; rho+pi as
funnel shifts; chi as one three-input bit operation per half; iota as an XOR
with the round constant) is evaluated symbolically: values known to be
constant are folded, and operations whose results never reach the read
output lanes are dropped.  For SHA3-256 of one 64-byte block (my_hhash,
my_hhash.h:28-35: the leaf chains, Merkle nodes and FRI leaves all hash
64 bytes) the state starts with 8 message lanes, the padding bytes 0x06 and
0x80 and 16 zero lanes, and only lanes 0..3 (the digest) are read.
"""
import math

import bench

RHO = [0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14]
RC = [0x0000000000000001, 0x0000000000008082, 0x800000000000808A, 0x8000000080008000, 0x000000000000808B,
      0x0000000080000001, 0x8000000080008081, 0x8000000000008009, 0x000000000000008A, 0x0000000000000088,
      0x0000000080008009, 0x000000008000000A, 0x000000008000808B, 0x800000000000008B, 0x8000000000008089,
      0x8000000000008003, 0x8000000000008002, 0x8000000000000080, 0x000000000000800A, 0x800000008000000A,
      0x8000000080008081, 0x8000000000008080, 0x0000000080000001, 0x8000000080008008]
M32 = 0xFFFFFFFF


class Graph:
    """Values are ("c", constant) or ("v", node index); a node holds its op
    count and the symbolic values it reads."""

    def __init__(self):
        self.nodes = []

    def node(self, cost, ins, cls="bit"):
        self.nodes.append((cost, [x for x in ins if x[0] == "v"], cls))
        return ("v", len(self.nodes) - 1)

    def xor(self, *ops):            # three-input XORs: ceil((m - 1) / 2) ops for m live inputs
        c, syms = 0, []
        for o in ops:
            if o[0] == "c":
                c ^= o[1]
            else:
                syms.append(o)
        m = len(syms) + (1 if c else 0)
        if not syms:
            return ("c", c)
        if m == 1:
            return syms[0]
        return self.node(math.ceil((m - 1) / 2), syms)

    def align(self, hi, lo, s):     # low 32 bits of (hi:lo) >> s: one funnel shift
        if hi[0] == "c" and lo[0] == "c":
            return ("c", (((hi[1] << 32) | lo[1]) >> s) & M32)
        return self.node(1, [hi, lo], "shift")

    def rot(self, lane, r):         # 64-bit rotate left of (lo, hi)
        lo, hi = lane
        r %= 64
        if r == 0:
            return lo, hi
        if r == 32:
            return hi, lo
        if r < 32:
            return self.align(lo, hi, 32 - r), self.align(hi, lo, 32 - r)
        r -= 32
        return self.align(hi, lo, 32 - r), self.align(lo, hi, 32 - r)

    def chi(self, a, b, c):         # a ^ (~b & c): one bit operation
        if b[0] == "c" and c[0] == "c":
            return self.xor(a, ("c", ~b[1] & c[1] & M32))
        if (c[0] == "c" and c[1] == 0) or (b[0] == "c" and b[1] == M32):
            return a
        if b[0] == "c" and b[1] == 0:
            return self.xor(a, c)
        return self.node(1, [a, b, c])

    def live_ops(self, outs, by_class=False):
        live, stack = set(), [o[1] for o in outs if o[0] == "v"]
        while stack:
            k = stack.pop()
            if k not in live:
                live.add(k)
                stack.extend(v[1] for v in self.nodes[k][1])
        if by_class:
            return {c: sum(self.nodes[k][0] for k in live if self.nodes[k][2] == c) for c in ("bit", "shift")}
        return sum(self.nodes[k][0] for k in live)


def keccak_ops(symbolic_lanes, consts, out_lanes, by_class=False):
    """Ops of one Keccak-f[1600] whose lanes < symbolic_lanes are unknown,
    lanes in consts are (lo, hi) constants, the rest zero; counted over the
    operations that reach lanes < out_lanes."""
    g = Graph()
    st = [(("c", 0), ("c", 0))] * 25
    for i in range(symbolic_lanes):
        st[i] = (g.node(0, []), g.node(0, []))
    for i, (lo, hi) in consts.items():
        st[i] = (("c", lo), ("c", hi))
    for rd in range(24):
        C = [(g.xor(*(st[x + 5 * y][0] for y in range(5))), g.xor(*(st[x + 5 * y][1] for y in range(5))))
             for x in range(5)]
        R = [g.rot(C[x], 1) for x in range(5)]
        A = [None] * 25
        for x in range(5):
            for y in range(5):
                i = x + 5 * y
                A[i] = tuple(g.xor(st[i][h], C[(x - 1) % 5][h], R[(x + 1) % 5][h]) for h in (0, 1))
        B = [None] * 25
        for x in range(5):
            for y in range(5):
                B[y + 5 * ((2 * x + 3 * y) % 5)] = g.rot(A[x + 5 * y], RHO[x + 5 * y])
        st = [tuple(g.chi(B[x + 5 * y][h], B[(x + 1) % 5 + 5 * y][h], B[(x + 2) % 5 + 5 * y][h]) for h in (0, 1))
              for y in range(5) for x in range(5)]
        st[0] = (g.xor(st[0][0], ("c", RC[rd] & M32)), g.xor(st[0][1], ("c", RC[rd] >> 32)))
    return g.live_ops([h for lane in st[:out_lanes] for h in lane], by_class)


def sha3_256_64_ops(by_class=False):
    # 64-byte message in lanes 0..7; 0x06 at byte 64 (lane 8), 0x80 at byte 135 (lane 16 high byte)
    return keccak_ops(8, {8: (0x06, 0), 16: (0, 0x80000000)}, 4, by_class)


def test_general_state_costs_the_full_round_count():
    """No constant lanes, every lane read: 24 rounds x (theta 80 + rho 48 +
    chi 50) plus iota's XORs (two per round, one where the round constant's
    high half is zero)."""
    zero_hi = sum(1 for rc in RC if rc >> 32 == 0)
    assert keccak_ops(25, {}, 25) == 24 * 180 - zero_hi


def test_bench_constant_is_the_derived_count():
    n = sha3_256_64_ops()
    assert n == bench.KECCAK_F_OPS
    assert 4100 < n < 24 * 180


def test_bench_constants_split_by_issue_class():
    """The roofline's peak prices the two classes apart: three-input bitwise
    operations (bitop3 / xor) and funnel shifts (alignbit)."""
    c = sha3_256_64_ops(by_class=True)
    assert c["bit"] == bench.KECCAK_F_BIT_OPS and c["shift"] == bench.KECCAK_F_SHIFT_OPS
    assert c["bit"] + c["shift"] == bench.KECCAK_F_OPS
    g = keccak_ops(25, {}, 25, by_class=True)
    assert g["shift"] == 24 * 58                 # theta's five rotations by one and rho's 24 lanes, two halves each
