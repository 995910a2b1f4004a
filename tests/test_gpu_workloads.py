"""GPU parity at the benchmark workloads' full sizes: the single 2^20 x 16
proof byte for byte against the oracle, the whole uneven batch of 256 proofs
against proofs made one at a time, and the proof server under Poisson
arrivals with a 2^20 x 24 request, each result against the same circuit
proved alone and checked by the verifiers."""
import threading
import time

import numpy as np
import pytest

import virgo_amd as V
from oracle import oracle as O
from virgo_amd import workloads as W

pytestmark = pytest.mark.gpu


def first_diff(a, b):
    return next((k for k in range(min(len(a), len(b))) if a[k] != b[k]), min(len(a), len(b)))


def test_c3_transcript_equals_oracle():
    """Random circuit 2^20 gates x 16 layers, 2^20 inputs, seed 1: every
    transcript byte equals the oracle's, the oracle's verifier accepts it and
    the output layer equals the oracle's evaluation."""
    circ = W.random_circuit(20, 16, seed=1)
    p = V.Prover(circ)
    tr = p.prove()
    print("gpu proof done, %d bytes; oracle proving" % len(tr), flush=True)
    ref, _ = O.prove(circ)
    assert len(tr) == len(ref)
    assert tr == ref, "transcripts differ first at byte %d" % first_diff(tr, ref)
    ok, msg = O.verify(circ, tr)
    assert ok, msg
    assert p.verify(tr)[0]
    assert np.array_equal(p.output(), O.evaluate(circ))
    p.close()


def test_wide_inputs_truncated_like_the_reference():
    """Inputs in [2^40, 2^63): both sides keep the low 32 bits as a signed
    int (zk_prover.cpp:57-66); transcripts equal, outputs equal."""
    for circ in (W.random_circuit(12, 4, seed=21, wide=True), W.random_circuit(17, 3, seed=22, wide=True)):
        assert int(np.count_nonzero(circ["u"][: circ["n"][0]] & 0x80000000)) > 0
        p = V.Prover(circ)
        tr = p.prove()
        ref, _ = O.prove(circ)
        assert tr == ref, "transcripts differ first at byte %d" % first_diff(tr, ref)
        assert np.array_equal(p.output(), O.evaluate(circ))
        p.close()


def test_c4_batch_of_256_equals_single_proofs():
    """The batched workload (256 proofs: matmul n log-uniform in {8..256},
    random layered log2 width U[10,18] x depth U[4,20], seed 7) through
    vg_prove_batch on 16 streams: every transcript equals the same circuit
    proved alone on a fresh context; the device verifier accepts every one;
    the largest random circuit and the largest matmul are accepted by the
    oracle's verifier, and the four smallest equal the oracle's proofs."""
    specs = W.batch_specs(256, seed=7)
    jobs = [(sp, W.make_job(sp)) for sp in specs]
    workers = 16
    streams = [V.Stream() for _ in range(workers)]
    provers = [V.Prover(c, stream=streams[k % workers]) for k, (_, c) in enumerate(jobs)]
    print("256 contexts, %.1f GB" % (sum(p.device_bytes() for p in provers) / 1e9), flush=True)
    out = V.prove_batch(provers, workers)
    again = V.prove_batch(provers, workers)
    assert again == out
    for k, (sp, c) in enumerate(jobs):
        single = V.prove(c)
        assert out[k] == single, "%s: batched differs from alone at byte %d" % (sp[0], first_diff(out[k], single))
        ok, why, _ = provers[k].verify(out[k])
        assert ok, (sp[0], why)
    print("batched == alone for all 256, all accepted", flush=True)

    def size(k):
        return int(jobs[k][1]["ty"].size)
    rnd = [k for k, (sp, _) in enumerate(jobs) if sp[1] == "random"]
    mm = [k for k, (sp, _) in enumerate(jobs) if sp[1] == "matmul"]
    for k in (max(rnd, key=size), max(mm, key=size)):
        print("oracle verify", specs[k][0], flush=True)
        ok, msg = O.verify(jobs[k][1], out[k])
        assert ok, (specs[k][0], msg)
    for k in sorted(range(len(jobs)), key=size)[:4]:
        assert out[k] == O.prove(jobs[k][1])[0], specs[k][0]
    for p in provers:
        p.close()
    for st in streams:
        st.close()


def test_c5_poisson_serving_equals_alone():
    """The serving workload (random layered log2 width U[10,20] x depth
    U[8,24], seed 11; 48 requests) plus one 2^20 x 24 circuit, submitted at
    Poisson arrival times (100 per second, seed 13) while earlier ones run:
    each transcript equals the circuit proved alone and is accepted by the
    device verifier; the largest is also accepted by the oracle's verifier."""
    jobs = W.serve_jobs(48)
    jobs.insert(20, ("random_w20_d24", W.random_circuit(20, 24, seed=77)))
    srv = V.Server(device_budget=96 << 30, slots=16)
    gaps = np.random.default_rng(13).exponential(1.0 / 100.0, len(jobs))
    ids = []

    def submit():
        t0 = time.perf_counter()
        arrive = t0 + np.cumsum(gaps)
        for k, (_, c) in enumerate(jobs):
            delay = arrive[k] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(c))

    t = threading.Thread(target=submit)
    t.start()
    t.join()
    res = []
    for jid, (name, _) in zip(ids, jobs):
        print("request", jid, name, flush=True)
        res.append(srv.wait(jid))
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert st["peak_admitted"] >= 2, st
    assert st["device_peak"] <= st["device_budget"]
    for (name, c), (tr, times) in zip(jobs, res):
        p = V.Prover(c)
        alone = p.prove()
        assert tr == alone, "%s: served differs from alone at byte %d" % (name, first_diff(tr, alone))
        ok, why, _ = p.verify(tr)
        assert ok, (name, why)
        assert times["latency"] >= times["prove"] > 0
        p.close()
    big = next(k for k, (name, _) in enumerate(jobs) if name == "random_w20_d24")
    print("oracle verify of the 2^20 x 24 request", flush=True)
    ok, msg = O.verify(jobs[big][1], res[big][0])
    assert ok, msg
