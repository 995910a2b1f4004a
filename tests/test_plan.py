"""Host planning through the C ABI (vg_plan): the circuit loader's checks,
padding and wiring counts, on the CPU (no GPU, no HIP call).

The loader checks every gate in one pass split into chunks of 2^16 gates on
the library's host threads; the first error in file order is the one
reported, as by the oracle's loader (zk_verifier::read_circuit and the gate
checks of zk_prover::evaluate, zk_verifier.cpp:15-217, zk_prover.cpp:54-157).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import virgo_amd as V
from virgo_amd import workloads as W
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def oracle_error(circ):
    with pytest.raises(RuntimeError) as e:
        O.prove(circ)
    return str(e.value).split(": ", 1)[1]


def plan_error(circ):
    with pytest.raises(V.VirgoError) as e:
        V.plan(circ)
    return str(e.value).split(": ", 1)[1]


@pytest.mark.parametrize("ty,u,v,msg", [(12, 0, 60, "more than 60"), (11, 0, 0, "unknown gate type"),
                                         (13, 1, 2, "u == v"), (5, 3, 3, "sum range"), (0, 0, 1 << 20, "out of bounds"),
                                         (-1, 0, 0, "unknown gate type")])
def test_plan_rejects_bad_gates_like_the_oracle(ty, u, v, msg):
    c = W.mixed_circuit(7, 2, seed=5)
    g = int(c["n"][0]) + 3
    c["ty"][g], c["u"][g], c["v"][g] = ty, u, v
    got = plan_error(c)
    assert msg in got
    assert got == oracle_error(c)


def wide(seed=3):
    """2^17 gates per layer (two loader chunks each), 4 layers."""
    return W.random_circuit(17, 4, seed=seed)


def layer_offset(c, layer):
    return int(np.sum(c["n"][:layer - 1]))


@pytest.mark.parametrize("plant", [
    # (file layer, gate, type, u, v) in planting order; the first in file order is reported
    [(4, 5, 11, 0, 0), (3, 100000, 0, 1 << 18, 0)],        # later layer first in the list
    [(3, 100000, 11, 0, 0), (3, 70000, 0, 0, 1 << 17)],     # same layer, second chunk before first's gate
    [(2, 131071, 13, 1, 2), (2, 65536, 5, 9, 9)],           # both in the second chunk
    [(2, 3, 0, -1, 0), (4, 0, 11, 0, 0)],
])
def test_plan_reports_the_first_error_in_file_order(plant):
    c = wide()
    for layer, g, ty, u, v in plant:
        k = layer_offset(c, layer) + g
        c["ty"][k], c["u"][k], c["v"][k] = ty, u, v
    got = plan_error(c)
    assert got == oracle_error(c)
    first = min(plant)
    assert ("layer %d gate %d" % first[:2]) in got or first[2] in (11, 13, -1)


def relay_circuit(bad_chunk2=False, later_bad_wire=False):
    """Inputs of 2^18, a layer of 2^18 add gates, then a layer of 2^17 gates
    (two loader chunks) whose gate 70000 is a relay reading u = 2^17 + 5: in
    range of the previous layer, beyond its own layer (zk_prover.cpp:496-497
    would index its beta table out of bounds)."""
    n0, n2 = 1 << 18, 1 << 17
    idx = np.arange(n0, dtype=np.int64)
    top = (np.zeros(n2, np.int32), np.arange(n2, dtype=np.int64), np.arange(n2, dtype=np.int64))
    top[0][70000], top[1][70000], top[2][70000] = 4, n2 + 5, 0
    if bad_chunk2:
        top[0][100000], top[1][100000], top[2][100000] = 4, n2 + 9, 0     # a second relay, later
    if later_bad_wire:
        top[1][120000] = n0                                              # a later wire past the previous layer
    return W._pack([(np.full(n0, 3, np.int32), idx % (1 << 20), np.zeros(n0, np.int64)),
                    (np.zeros(n0, np.int32), idx, idx), top])


def test_plan_relay_checked_after_the_layers_other_gates():
    """The library's own refusal (the reference reads past a table there;
    the oracle restates the reads and has no such check): the first relay
    gate of the layer, and only when the layer has no other bad gate."""
    got = plan_error(relay_circuit())
    assert got.startswith("layer 3: relay gate 70000 reads u=131077 beyond")
    assert plan_error(relay_circuit(bad_chunk2=True)) == got
    c = relay_circuit(later_bad_wire=True)
    got = plan_error(c)
    assert got == "layer 3 gate 120000: input index out of bounds" and got == oracle_error(c)


def test_plan_of_good_circuits():
    for c in (W.matmul_circuit(8, seed=1), W.random_circuit(12, 6, seed=2), W.mixed_circuit(10, 5, seed=4)):
        dev, pin = V.plan(c)
        assert dev > 0 and pin > 0
        assert V.plan(c) == (dev, pin)


PROBE = r"""
import sys, json
sys.path.insert(0, %r)
import virgo_amd as V
from virgo_amd import workloads as W
cs = [W.random_circuit(17, 4, seed=3), W.mixed_circuit(12, 5, seed=7), W.matmul_circuit(32, seed=1),
      W.random_circuit(18, 3, seed=9)]
print(json.dumps([V.plan(c) for c in cs]))
"""


def test_plan_independent_of_host_threads():
    """The chunked pass merges per-chunk counts (records by u and by v, the
    phase-2 zero list): one thread and eight give the same plan."""
    outs = []
    for th in ("1", "8"):
        env = dict(os.environ, VIRGO_HOST_THREADS=th)
        r = subprocess.run([sys.executable, "-c", PROBE % ROOT], env=env, capture_output=True, text=True, check=True)
        outs.append(r.stdout.strip())
    assert outs[0] == outs[1]


def test_plan_from_many_threads_at_once():
    """Several callers plan at the same time (a proof server's workers): they
    share the library's host pool and recycled blocks; every plan equals the
    one made alone."""
    import threading
    cs = [W.random_circuit(17, 3, seed=s) for s in range(3)] + [W.mixed_circuit(11, 4, seed=1), W.matmul_circuit(16, seed=3)]
    want = [V.plan(c) for c in cs]
    got, errs = {}, []

    def work(t):
        try:
            for r in range(4):
                k = (t + r) % len(cs)
                got[(t, r)] = (k, V.plan(cs[k]))
        except Exception as e:          # surfaced below
            errs.append(e)
    th = [threading.Thread(target=work, args=(t,)) for t in range(8)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errs
    assert len(got) == 32 and all(v == want[k] for k, v in got.values())
