"""GPU: the proof server (continuous batching, vg_server_*).  Every request's
transcript must be byte-identical to the same circuit proved alone, whatever
else is in flight, and the HBM budget must bound what is admitted at once."""
import threading
import time

import pytest

import virgo_amd as V
from oracle import oracle as O
from virgo_amd import workloads as W

pytestmark = pytest.mark.gpu

MIX = [
    ("matmul8", lambda: W.matmul_circuit(8, seed=2)),
    ("random_w10_d4", lambda: W.random_circuit(10, 4, seed=3)),
    ("random_w12_d3", lambda: W.random_circuit(12, 3, seed=4)),
    ("matmul16", lambda: W.matmul_circuit(16, seed=5)),
    ("random_w14_d6", lambda: W.random_circuit(14, 6, seed=6)),
    ("random_w11_d9", lambda: W.random_circuit(11, 9, seed=7)),
    ("random_w16_d4", lambda: W.random_circuit(16, 4, seed=8)),
    ("random_w13_d3_skew", lambda: W.random_circuit(13, 3, seed=9, v_bits=6)),
    ("mixed_w12_d4", lambda: W.mixed_circuit(12, 4, seed=10)),
]


@pytest.fixture(scope="module")
def alone():
    circs = {name: f() for name, f in MIX}
    return circs, {name: V.prove(c) for name, c in circs.items()}


def test_server_mix_matches_alone(alone):
    """Requests arrive while others run (staggered submissions from a second
    thread); each result equals the proof made alone, and the two smallest
    equal the oracle's."""
    circs, ref = alone
    srv = V.Server(device_budget=8 << 30, slots=6)
    order = [name for name, _ in MIX] * 2
    ids = {}

    def submit():
        for k, name in enumerate(order):
            ids[k] = srv.submit(circs[name])
            time.sleep(0.002 * (k % 3))

    t = threading.Thread(target=submit)
    t.start()
    t.join()
    for k, name in enumerate(order):
        print("request", k, name, flush=True)
        tr, times = srv.wait(ids[k])
        assert tr == ref[name], name
        assert times["latency"] >= times["prove"] > 0
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(order) and st["failed"] == 0
    assert st["peak_admitted"] >= 2, st          # proofs overlapped on the device
    assert st["device_peak"] <= st["device_budget"]
    for name in ("matmul8", "random_w10_d4"):
        want, _ = O.prove(circs[name])
        assert ref[name] == want, name


def test_server_budget_bounds_admission(alone):
    """A budget that holds only the largest lease once: requests still all
    finish, bit-exact, never exceeding the budget; a request larger than the
    budget fails alone and the rest are unaffected."""
    circs, ref = alone
    big = V.Prover(circs["random_w16_d4"])
    need = big.device_bytes()
    big.close()
    srv = V.Server(device_budget=need + (64 << 20), slots=4)
    names = ["random_w16_d4", "matmul8", "random_w14_d6", "random_w16_d4", "random_w10_d4"]
    ids = [srv.submit(circs[n]) for n in names]
    for jid, n in zip(ids, names):
        print("request", jid, n, flush=True)
        tr, _ = srv.wait(jid)
        assert tr == ref[n], n
    st = srv.stats()
    assert st["device_peak"] <= st["device_budget"]
    assert st["failed"] == 0
    # a circuit whose lease exceeds the whole budget
    huge = W.random_circuit(18, 4, seed=11)
    jid = srv.submit(huge)
    ok = srv.submit(circs["matmul8"])
    with pytest.raises(V.VirgoError, match="server holds"):
        srv.wait(jid)
    tr, _ = srv.wait(ok)
    assert tr == ref["matmul8"]
    srv.close()


def test_server_repeat_and_collect_errors(alone):
    circs, ref = alone
    srv = V.Server(device_budget=2 << 30, slots=2)
    ids = [srv.submit(circs["random_w12_d3"]) for _ in range(5)]
    for jid in ids:
        assert srv.wait(jid)[0] == ref["random_w12_d3"]
    with pytest.raises(V.VirgoError, match="unknown or already collected"):
        srv.wait(ids[0])
    srv.close()


def test_server_bad_request_fails_alone(alone):
    """A request the circuit loader rejects fails on its own wait; the
    requests around it are unaffected; a server needs a budget."""
    circs, ref = alone
    with pytest.raises(V.VirgoError, match="bad budget"):
        V.Server(device_budget=0, slots=2)
    srv = V.Server(device_budget=2 << 30, slots=2)
    bad = dict(circs["random_w10_d4"])
    bad["u"] = bad["u"].copy()
    bad["u"][-1] = 1 << 20                      # past the previous layer
    jb = srv.submit(bad)
    jg = srv.submit(circs["matmul8"])
    with pytest.raises(V.VirgoError, match="out of bounds"):
        srv.wait(jb)
    assert srv.wait(jg)[0] == ref["matmul8"]
    st = srv.stats()
    srv.close()
    assert st["failed"] == 1 and st["completed"] == 2


def test_plan_sizes_equal_the_contexts(alone):
    """vg_plan (host only) reports what a context of the circuit on a
    caller's stream allocates: the device bytes up to the 256-byte alignment
    of its two blocks."""
    circs, _ = alone
    st = V.Stream()
    try:
        for name, c in circs.items():
            dev, pin = V.plan(c)
            p = V.Prover(c, stream=st)
            held = p.device_bytes()
            p.close()
            assert 0 <= dev - held < 512, name
            assert pin > 0
    finally:
        st.close()
