"""Device-accelerated verifier (vg_verify) against the oracle's verifier.

The oracle restates zk_verifier::verify and verify_poly_commitment on the CPU
(oracle/vg_oracle.c, vgo_verify); vg_verify must reach the same verdict on
every transcript: accept each GPU proof, and agree on transcripts with one
byte changed anywhere (most changes are rejected; a few touch bytes the
reference verifier never reads, such as the odd entries of the final
codewords, and both accept those) or truncated / extended."""
import numpy as np
import pytest

import virgo_amd as V
from oracle import oracle as O
from virgo_amd import workloads as W

CASES = {
    "matmul8": lambda: W.matmul_circuit(8, seed=2),
    "matmul32": lambda: W.matmul_circuit(32, seed=1),
    "random_w10_d4": lambda: W.random_circuit(10, 4, seed=3),
    "random_w13_d3_skew": lambda: W.random_circuit(13, 3, seed=6, v_bits=6),
    # the device predicates of every gate type a description may use
    "mixed_w10_d5": lambda: W.mixed_circuit(10, 5, seed=12),
    "mixed_w14_d3": lambda: W.mixed_circuit(14, 3, seed=17),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_verify_accepts_gpu_proof(name):
    circ = CASES[name]()
    print("case", name, flush=True)
    p = V.Prover(circ)
    try:
        tr = p.prove()
        ok, why, ms = p.verify(tr)
        assert ok, why
        ok_o, why_o = O.verify(circ, tr)
        assert ok_o, why_o
        # verifying twice, and after another proof, gives the same verdict
        assert p.verify(tr)[0]
        assert p.prove() == tr
        assert p.verify(tr)[0]
    finally:
        p.close()


def _flip(tr, pos, bit=0):
    b = bytearray(tr)
    b[pos] ^= 1 << bit
    return bytes(b)


@pytest.mark.gpu
def test_verify_tampered_matches_oracle():
    circ = W.random_circuit(12, 3, seed=4)
    p = V.Prover(circ)
    try:
        tr = p.prove()
        n = len(tr)
        rng = np.random.default_rng(5)
        # every region: header, GKR messages, roots and sums, FRI roots, final
        # codewords, query answers and paths, the fft_gkr section
        pos = sorted(set([0, 5, 13, 40, 200, 1000, n // 20, n // 10, n // 5, n // 3, n // 2, 2 * n // 3,
                          n - 5000, n - 500, n - 20, n - 1] + [int(x) for x in rng.integers(0, n, 40)]))
        rejected = 0
        for q in pos:
            t = _flip(tr, q, q % 8)
            ok, why, _ = p.verify(t)
            ok_o, why_o = O.verify(circ, t)
            assert ok == ok_o, "byte %d of %d: device %s (%s), oracle %s (%s)" % (q, n, ok, why, ok_o, why_o)
            rejected += not ok
        assert rejected >= len(pos) - 4
        # V_res + p: congruent to the honest value but not canonical
        depth = int.from_bytes(tr[4:8], "little")
        at = 12 + 4 * depth + 8
        b = bytearray(tr)
        b[at:at + 8] = (int.from_bytes(tr[at:at + 8], "little") + (1 << 61) - 1).to_bytes(8, "little")
        ok, why, _ = p.verify(bytes(b))
        assert not ok and "canonical" in why
        assert not O.verify(circ, bytes(b))[0]
        for t in (tr[:-1], tr[: n // 2], tr + b"\0"):
            ok, why, _ = p.verify(t)
            ok_o, _ = O.verify(circ, t)
            assert not ok and not ok_o, why
        assert p.verify(tr)[0]
    finally:
        p.close()


@pytest.mark.gpu
def test_verify_mixed_rejects_changed_wiring():
    """A power-sum range one entry shorter, a not gate turned into a relay,
    or a bit test on another wire: the device verifier rejects the proof of
    the original circuit, like the oracle's."""
    circ = W.mixed_circuit(11, 3, seed=18)
    p = V.Prover(circ)
    tr = p.prove()
    p.close()
    base = int(circ["n"][0])
    lay = slice(base, base + int(circ["n"][1]))
    ty, u, v = circ["ty"][lay], circ["u"][lay], circ["v"][lay]
    g12 = base + int(np.flatnonzero((ty == 12) & (v > u))[0])
    g6 = base + int(np.flatnonzero(ty == 6)[0])
    g13 = base + int(np.flatnonzero(ty == 13)[0])
    for what, g, field, val in (("power sum", g12, "v", circ["v"][g12] - 1), ("not", g6, "ty", 10),
                                ("bit test", g13, "uv", (circ["u"][g13] + 1) % int(circ["n"][0]))):
        c2 = {k: a.copy() for k, a in circ.items()}
        if field == "uv":
            c2["u"][g] = c2["v"][g] = val
        else:
            c2[field][g] = val
        print("changed", what, flush=True)
        q = V.Prover(c2)
        try:
            ok, why, _ = q.verify(tr)
        finally:
            q.close()
        assert not ok, what
        assert not O.verify(c2, tr)[0], what


@pytest.mark.gpu
def test_verify_rejects_other_circuits_proof():
    a, b = W.random_circuit(11, 4, seed=21), W.random_circuit(11, 4, seed=22)
    pa, pb = V.Prover(a), V.Prover(b)
    try:
        ta = pa.prove()
        ok, why, _ = pb.verify(ta)
        assert not ok
        assert not O.verify(b, ta)[0]
    finally:
        pa.close()
        pb.close()


@pytest.mark.gpu
def test_verify_full_size_c3():
    """2^20 gates x 16 layers: the device verifier accepts the GPU proof and
    rejects it with one value of a query answer changed."""
    circ = W.random_circuit(20, 16, seed=1)
    p = V.Prover(circ)
    try:
        tr = p.prove()
        ok, why, ms = p.verify(tr)
        print("C3 verify ms", round(ms, 1), flush=True)
        assert ok, why
        ok, why, _ = p.verify(_flip(tr, len(tr) - 40000))
        assert not ok
    finally:
        p.close()
